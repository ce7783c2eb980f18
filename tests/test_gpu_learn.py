"""The learned pressure warm start on the GPU: the network's forward pass and
its parameter gradients against the CPU reference (oracle/unet_ref.py), and
the warm start inside the pressure solve.

Tolerances: the GPU stores weights and activations in bfloat16 and
accumulates in FP32; the reference run with the same bfloat16 roundings
(bf16=True) differs only where an FP32 sum lands on the other side of a
bfloat16 rounding boundary, so outputs agree to 2e-2 of their range and
gradients to 3e-2 (relative L2 per layer). Against the pure FP64 reference
(no rounding) the gradient directions agree to cosine 0.98 per layer."""
import numpy as np
import pytest

import unet_ref as U
from gpu_scenes import watersim_hip

pytestmark = pytest.mark.gpu


def scene(n, seed):
    rng = np.random.default_rng(seed)
    types = rng.choice([0, 1, 2], size=(n, n, n), p=[0.3, 0.5, 0.2])
    r0 = rng.normal(size=(n, n, n))
    q = rng.normal(size=(n, n, n))
    x = np.zeros((n, n, n, 5))
    fl = types == 1
    x[..., 0] = np.where(fl, r0, 0)
    x[..., 1] = types == 0
    x[..., 2] = fl
    x[..., 3] = types == 2
    x[..., 4] = np.where(fl, q, 0)
    return x, fl, rng


def params(seed):
    rng = np.random.default_rng(seed)
    th = U.init_params(seed)
    return th + rng.normal(0, 0.02, size=th.shape)


def solver():
    return watersim_hip.Solver(16, 16, 16, 0.1, 16)


def test_param_count_matches_reference():
    s = solver()
    assert len(s.net_params()) == U.param_count()
    s.close()


@pytest.mark.parametrize("n", [8, 16])
def test_forward_matches_reference(n):
    x, fl, _ = scene(n, 11)
    th = params(7)
    s = solver()
    s.set_net_params(th.astype(np.float32))
    got = s.net_forward(x)
    want, _ = U.forward(th.astype(np.float32).astype(np.float64), x, bf=True)
    scale = np.abs(want).max()
    err = np.abs(got - want).max() / scale
    mean = np.abs(got - want).mean() / scale
    print("  forward %d^3: max rel err %.2e, mean %.2e (scale %.3g)" % (n, err, mean, scale))
    assert err <= 2e-2 and mean <= 2e-3
    s.close()


def test_gradients_match_reference():
    n = 16
    x, fl, rng = scene(n, 12)
    tgt = rng.normal(size=(n, n, n))
    th = params(8).astype(np.float32).astype(np.float64)
    s = solver()
    s.set_net_params(th.astype(np.float32))
    loss_g, g = s.net_grad(x, tgt, fl)
    out, cache = U.forward(th, x, bf=True)
    loss_r, gout = U.loss(out, tgt, fl)
    gr = U.backward(th, cache, gout, bf=True)
    out64, cache64 = U.forward(th, x)
    _, gout64 = U.loss(out64, tgt, fl)
    g64 = U.backward(th, cache64, gout64)
    print("  loss gpu %.6f ref %.6f" % (loss_g, loss_r))
    assert abs(loss_g - loss_r) <= 1e-2 * loss_r
    off = 0
    for (name, ci, co, k, _), (_, nw, nb) in zip(U.LAYERS, U.layer_sizes()):
        sl = slice(off, off + nw + nb)
        a, b, c = g[sl].astype(np.float64), gr[sl], g64[sl]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = float(a @ c / max(np.linalg.norm(a) * np.linalg.norm(c), 1e-30))
        print("  %-5s grad rel L2 vs bf16 ref %.2e, cosine vs fp64 %.5f" % (name, rel, cos))
        assert rel <= 3e-2, name
        assert cos >= 0.98, name
        off += nw + nb
    s.close()


def test_warm_start_runs_inside_the_pressure_solve():
    """With the warm start on and zero weights the prediction is zero, so the
    solve must equal the plain solve bit for bit; with trained-looking random
    weights it must still converge to the same residual."""
    import golden
    from gpu_scenes import gpu_from_golden
    g = golden.load("cfg1_32")
    qs = []
    for ws in (0, 1):
        s = gpu_from_golden(g)
        s.set_net_params(np.zeros(U.param_count(), np.float32))
        s.set(warmStart=ws, pressureTol=1e-4, maxPressureIters=2000, residualEvery=8)
        s.stage("rasterize")
        for k in "uvw":
            s.upload(k, g["grav_" + k])
        s.upload("q", g["in_q"])
        s.stage("pressure")
        qs.append(s.download("q"))
        s.close()
    assert np.array_equal(qs[0], qs[1])


def test_tiled_and_untiled_convolutions_agree():
    """The LDS-staged convolution and the untiled one compute the same
    products (different FP32 summation order only)."""
    x, fl, rng = scene(16, 13)
    tgt = rng.normal(size=(16, 16, 16))
    th = params(9).astype(np.float32)
    res = []
    for tiled in (1, 0):
        s = solver()
        s.set(netTiled=tiled)
        s.set_net_params(th)
        out = s.net_forward(x)
        _, g = s.net_grad(x, tgt, fl)
        res.append((out, g))
        s.close()
    (o1, g1), (o0, g0) = res
    eo = np.abs(o1 - o0).max() / np.abs(o0).max()
    eg = np.linalg.norm(g1 - g0) / np.linalg.norm(g0)
    print("  tiled vs untiled: out %.2e, grad %.2e" % (eo, eg))
    assert eo <= 1e-2 and eg <= 1e-2
