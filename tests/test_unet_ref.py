"""Pins the CPU reference of the warm-start network (oracle/unet_ref.py),
which has no counterpart in the reference solver, by mathematics:
  - a 3x3x3 convolution against a brute-force loop over voxels and taps;
  - pooling / upsampling against their definitions, and as adjoints;
  - the full backward pass against central finite differences in FP64;
  - bf16 rounding against the known bit patterns.
A dropped tap, a flipped offset, a wrong concat order or a missing skip
gradient fails one of these."""
import numpy as np

import unet_ref as U


def test_param_count():
    assert U.param_count() == 355009


def test_conv_matches_brute_force():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(4, 5, 6, 3))
    W = rng.normal(size=(2, 27, 3))
    b = rng.normal(size=2)
    y = U.conv(x, W, b, 3)
    Z, Y, X, _ = x.shape
    for z in range(Z):
        for yy in range(Y):
            for xx in range(X):
                acc = b.copy()
                for t, (dz, dy, dx) in enumerate(U.OFFS):
                    zz, y2, x2 = z + dz, yy + dy, xx + dx
                    if 0 <= zz < Z and 0 <= y2 < Y and 0 <= x2 < X:
                        acc += W[:, t, :] @ x[zz, y2, x2]
                assert np.allclose(y[z, yy, xx], acc)


def test_pool_up_and_their_adjoints():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(4, 4, 4, 2))
    p = U.pool(x)
    assert np.isclose(p[1, 0, 1, 1], x[2:4, 0:2, 2:4, 1].mean())
    u = U.up(p)
    assert np.array_equal(u[3, 1, 2], p[1, 0, 1])
    g = rng.normal(size=p.shape)
    # <pool(x), g> = <x, pool_back(g)>, <up(p), h> = <p, up_back(h)>
    assert np.isclose((U.pool(x) * g).sum(), (x * U.pool_back(g)).sum())
    h = rng.normal(size=x.shape)
    assert np.isclose((U.up(p) * h).sum(), (p * U.up_back(h)).sum())


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    theta = U.init_params(5)
    theta += rng.normal(0, 0.01, size=theta.shape)  # non-zero biases too
    inp = rng.normal(size=(8, 8, 8, 5))
    tgt = rng.normal(size=(8, 8, 8))
    mask = rng.random((8, 8, 8)) < 0.6

    def f(th):
        out, _ = U.forward(th, inp)
        return U.loss(out, tgt, mask)[0]

    out, cache = U.forward(theta, inp)
    L, gout = U.loss(out, tgt, mask)
    g = U.backward(theta, cache, gout)
    # a spread of parameters over every layer (weights and biases)
    sizes = U.layer_sizes()
    idx, off = [], 0
    for _, nw, nb in sizes:
        idx += list(rng.choice(nw, 3, replace=False) + off)
        idx.append(off + nw + rng.integers(nb))
        off += nw + nb
    h = 1e-6
    for i in idx:
        e = np.zeros_like(theta)
        e[i] = h
        fd = (f(theta + e) - f(theta - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.140625, 1e-3])
    r = U.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie at 1 + 2^-8 rounds to even (1.0)
    assert r[2] == 1.0 + 2 ** -7            # above the tie rounds up
    assert r[3] == -2.5
    assert r[4] == 3.140625
    assert abs(r[5] - 1e-3) <= 1e-3 * 2 ** -8
