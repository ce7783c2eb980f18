#!/usr/bin/env python3
"""Trains the learned pressure warm start on the GPU.

Pairs come from plain runs of the planet substep: for each training seed a
world is built, the solver runs `--warmup` substeps, then `--record`
substeps whose pressure solves go to a tight residual (`--tight`) with
recording on, so every solve stores (inputs, converged correction) in the
device sample bank. With --coarse-grid 1 (the default) the network sees the
grid coarsened 2x per axis and learns the block mean of the correction. Training (Adam, random crops of the recorded grids) is
interleaved with recording. The weights are written as raw float32 in the
flat order of include/watersim_learn.h.

    python3 watersim-v1_amd/train_warmstart.py --out watersim-v1_amd/weights/warmstart_unet.f32
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import watersim_hip  # noqa: E402

W_DOMAIN = 5.1


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--target", type=int, default=2117000)
    ap.add_argument("--seeds", type=int, default=16)
    ap.add_argument("--seed0", type=int, default=20261000, help="first training world seed")
    ap.add_argument("--warmup", type=int, default=24)
    ap.add_argument("--record", type=int, default=96)
    ap.add_argument("--tight", type=float, default=1e-5)
    ap.add_argument("--bank", type=int, default=384)
    ap.add_argument("--crop", type=int, default=64)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--train-every", type=int, default=24, help="recorded substeps between training bursts")
    ap.add_argument("--burst", type=int, default=24, help="updates per training burst")
    ap.add_argument("--final", type=int, default=600, help="updates after the last seed")
    ap.add_argument("--coarse", type=float, default=1.0,
                    help="weight of the loss term on the 4x4x4-averaged error (the smooth modes SOR is slow on)")
    ap.add_argument("--coarse-grid", type=int, default=1,
                    help="1: the network sees the grid coarsened 2x per axis and predicts a correction constant "
                         "over 2x2x2 blocks (netCoarse)")
    ap.add_argument("--init", default=None, help="start from these weights")
    ap.add_argument("--out", required=True)
    a = ap.parse_args()

    t0 = time.time()
    history = []
    pairs = 0
    s = None
    for k in range(a.seeds):
        seed = a.seed0 + k
        dx = W_DOMAIN / a.n
        world = watersim_hip.build_world(a.n, dx, 0.25 * W_DOMAIN, 0.15 * W_DOMAIN, seed, a.target)
        if s is None:
            # one context for every world: the sample bank and the optimiser
            # state carry over from seed to seed
            s = watersim_hip.Solver(a.n, a.n, a.n, dx, a.target + 8192)
            w0 = np.fromfile(a.init, dtype=np.float32) if a.init else init_weights(
                np.random.default_rng(a.seed0), len(s.net_params()))
            s.set_net_params(w0)
            s.set(netCoarse=a.coarse_grid)
            s.net_bank(a.bank)
            s.set(netLossCoarse=a.coarse)
        watersim_hip.load_world(s, world)
        s.set(pic=0.05, omega=1.7, gravity=9.8, maxSpeed=3.2, vorticity=0.05, cfl=3.0)
        s.set(pressureTol=1e-4, maxPressureIters=4000, residualEvery=8, recordSamples=0)
        s.run_substeps(a.warmup)
        s.set(pressureTol=a.tight, maxPressureIters=3000, recordSamples=1)
        done = 0
        while done < a.record:
            m = min(a.train_every, a.record - done)
            s.run_substeps(m)
            done += m
            pairs += m
            history.append(s.net_train(a.burst, a.batch, a.crop, a.lr, seed=seed * 131 + done))
            log("  seed %d: %d recorded, loss %.4f, %.0f s" % (seed, done, history[-1], time.time() - t0))
        s.set(recordSamples=0, pressureTol=1e-4)
        log("seed %d: %d pairs so far, bank %d, loss %.4f, %.0f s" % (seed, pairs, s.net_bank_count(),
                                                                    history[-1], time.time() - t0))
    # a final pass over the last world's bank at a lower rate
    for i in range(0, a.final, 50):
        history.append(s.net_train(50, a.batch, a.crop, a.lr * 0.3, seed=7 + i))
    w = s.net_params()
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    w.astype(np.float32).tofile(a.out)
    meta = {"pairs": pairs, "net_coarse": a.coarse_grid, "seeds": [a.seed0 + k for k in range(a.seeds)],
            "grid": a.n, "crop": a.crop, "bank": a.bank, "train_every": a.train_every, "burst": a.burst,
            "final": a.final, "record_per_seed": a.record, "warmup": a.warmup,
            "batch": a.batch, "lr": a.lr, "tight_residual": a.tight, "loss_history": history,
            "seconds": time.time() - t0}
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    s.close()
    log("wrote %s (%d params), final loss %.4f, %.0f s" % (a.out, len(w), history[-1], time.time() - t0))


LAYERS = [(5, 16, 27), (16, 16, 27), (16, 32, 27), (32, 32, 27), (32, 64, 27), (64, 64, 27), (96, 32, 27),
          (32, 32, 27), (48, 16, 27), (16, 16, 27), (16, 1, 1)]


def init_weights(rng, n):
    """He-normal weights for the ReLU layers, small weights for the linear
    head, zero biases, in the flat order W[cout][tap][cin], b[cout]."""
    parts = []
    for i, (ci, co, taps) in enumerate(LAYERS):
        fan = ci * taps
        std = np.sqrt(2.0 / fan) if i < len(LAYERS) - 1 else 0.1 / np.sqrt(fan)
        parts.append(rng.normal(0, std, size=co * taps * ci))
        parts.append(np.zeros(co))
    w = np.concatenate(parts).astype(np.float32)
    assert len(w) == n, (len(w), n)
    return w


if __name__ == "__main__":
    main()
