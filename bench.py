#!/usr/bin/env python3
"""Benchmark: FLIP substeps/s of the planet substep pipeline on one MI355X.

Workload (default "cfg3"): 128^3 MAC grid, ~2.0M FLIP particles, seeded
voxel planet (core 0.25 W, sea level 0.4 W, W = 5.1 m), FLIP 0.95,
red-black SOR (omega 1.7) to relative residual 1e-4, vorticity confinement
0.05, dt from the CFL law. Synthetic input built on the host by the
library's world builder; all state is resident in HBM before timing.

A "step" is one substep. Prints one JSON line (diagnostics on stderr).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "watersim-v1_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

W_DOMAIN = 5.1
CONFIGS = {
    # name: grid n, particle target, seed, vorticity, SOR mode
    "cfg1": dict(n=32, target=10000, seed=20260114, vorticity=0.0, tol=0.0, iters=40),
    "cfg2": dict(n=64, target=232000, seed=20260115, vorticity=0.05, tol=1e-4, iters=0),
    "cfg3": dict(n=128, target=2117000, seed=20260116, vorticity=0.05, tol=1e-4, iters=0),
}
HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_solver(cfg):
    import watersim_hip
    n = cfg["n"]
    dx = W_DOMAIN / n
    t = time.time()
    world = watersim_hip.build_world(n, dx, 0.25 * W_DOMAIN, 0.15 * W_DOMAIN, cfg["seed"], cfg["target"])
    log("world %d^3: %d particles, land %.3f, built in %.1fs" % (n, world["nP"], world["landFrac"], time.time() - t))
    s = watersim_hip.solver_for_world(world)
    s.set(pic=0.05, omega=1.7, gravity=9.8, maxSpeed=3.2, vorticity=cfg["vorticity"], cfl=3.0)
    if cfg["tol"] > 0:
        s.set(pressureTol=cfg["tol"], maxPressureIters=4000, residualEvery=8)
    else:
        s.set(pressureIters=cfg["iters"])
    return s, world


def cpu_baseline(s, world, cfg, budget_s=12.0, max_substeps=6):
    """The oracle (single-threaded C port of the reference substep, reference
    lexicographic SOR order) on the same state, for a bounded sample."""
    import cpu_oracle as oracle
    n = cfg["n"]
    st = oracle.OracleState({"nx": n, "ny": n, "nz": n, "dx": W_DOMAIN / n, "coreR": world["coreR"],
                             "oceanR": world["oceanR"], "spacing": world["spacing"],
                             "particleVolume": world["particleVolume"], "dv": world["dv"], "tn": world["tn"]})
    st.set_static(world["solid"], world["R"], world["staticTypes"], world["gu"], world["gv"], world["gw"])
    nP = int(s.get("nP"))
    parts = {k: s.download(k) for k in ["px", "py", "pz", "pvx", "pvy", "pvz", "pT", "pAir", "pflag", "pcool"]}
    st.set_particles(nP, **parts)
    st.arrays["q"][:] = s.download("q")
    st.set_params(pic=0.05, omega=1.7, gravity=9.8, maxSpeed=3.2, vorticity=cfg["vorticity"], cfl=3.0,
                  gRamp=s.get("gRamp"), umax=s.get("umax"))
    if cfg["tol"] > 0:
        st.set_params(pressureTol=cfg["tol"], maxPressureIters=4000, residualEvery=8, sorRedBlack=0)
    else:
        st.set_params(pressureIters=cfg["iters"])
    done, t0 = 0, time.perf_counter()
    while done < max_substeps and time.perf_counter() - t0 < budget_s:
        st.call("oc_substep", st.call("oc_dt_law"))
        done += 1
    el = time.perf_counter() - t0
    return {"value": done / el, "unit": "substeps/s", "cores": 1, "kind": "port",
            "sample": "%d substeps of the same %d^3 / %d-particle state (oracle: reference arithmetic, "
                      "lexicographic SOR to %g), single thread" % (done, n, nP, cfg["tol"] or cfg["iters"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="cfg3", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the planet substep runs on a single MI355X; "
              "--gpus %d is not supported (see DESIGN.md for how it would shard)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    cfg = CONFIGS[a.config]
    s, world = make_solver(cfg)
    n, nP = cfg["n"], int(s.get("nP"))
    log("warmup %d substeps" % a.warmup)
    for i in range(a.warmup):
        s.run_substeps(1)
    s.sync()
    s.set(timing=1)
    log("timing %d substeps" % a.steps)
    s.sync()
    t0 = time.perf_counter()
    s.run_substeps(a.steps)
    s.sync()
    el = time.perf_counter() - t0
    sor_ms, sor_n = s.get("sorMs"), s.get("sorLaunches")
    p_ms, p_n, sweeps = s.get("pressureMs"), s.get("pressureSolves"), s.get("sweepsTotal")
    s.set(timing=0)
    sps = a.steps / el
    C = n ** 3
    # dominant kernel: one colour of a red-black SOR sweep; algorithmic bytes
    # 13 B per cell per full sweep (q read+write, rhs, type), half per launch
    sor_avg_s = (sor_ms / 1e3) / max(sor_n, 1)
    sor_bytes = 6.5 * C
    achieved = sor_bytes / sor_avg_s / 1e9 if sor_n else 0.0
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "sor_traffic_%s.json" % a.config)
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")
    ke = s.kinetic_energy()
    log("substep %.3f ms, pressure %.3f ms/solve, %.1f sweeps/substep, SOR launch %.2f us, ke %.4g"
        % (1e3 * el / a.steps, p_ms / max(p_n, 1), sweeps / a.steps, 1e6 * sor_avg_s, ke))
    out = {
        "metric": "FLIP substeps/s at 128^3, 2M particles",
        "value": sps,
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * el / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded planet built by the library's world builder)",
        "config": {"workload": "%s: %d^3 MAC grid, %d FLIP particles, SOR %s, vorticity %g; "
                               "thermal tick not included" % (a.config, n, nP,
                                                              "to residual %g" % cfg["tol"] if cfg["tol"]
                                                              else "%d sweeps" % cfg["iters"], cfg["vorticity"]),
                   "grid": n, "particles": nP, "particle_updates_per_s": sps * nP,
                   "pressure_ms_per_solve": p_ms / max(p_n, 1), "sor_sweeps_per_substep": sweeps / a.steps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / (HBM_PEAK / 1e9), "traffic": traffic,
                     "kernel": "k_sor (one colour of a red-black sweep)",
                     "avg_launch_us": 1e6 * sor_avg_s, "algorithmic_bytes_per_launch": sor_bytes},
    }
    if not a.no_cpu_baseline:
        log("cpu baseline (oracle, 1 thread)")
        out["cpu_baseline"] = cpu_baseline(s, world, cfg)
    s.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
