"""Synthetic inputs the benchmark adds beyond the reference: the
divergence-free curl-noise initial velocity and the Kelvin temperature map."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_curl_noise_is_divergence_free_with_the_requested_rms():
    rng = np.random.default_rng(3)
    p = rng.uniform(0.5, 4.6, size=(3, 400))
    h = 1e-4
    pts = [p]
    for a in range(3):
        for sgn in (1, -1):
            q = p.copy()
            q[a] += sgn * h
            pts.append(q)
    allp = np.concatenate(pts, axis=1)  # one call: one normalisation for every point
    u = bench.curl_noise_velocity(allp[0], allp[1], allp[2], 20260115, 0.5, 5.1).astype(np.float64)
    n = p.shape[1]
    blk = [u[:, i * n:(i + 1) * n] for i in range(7)]
    div = sum((blk[1 + 2 * a][a] - blk[2 + 2 * a][a]) / (2 * h) for a in range(3))
    grad_scale = np.abs(blk[1] - blk[2]).max() / (2 * h)
    assert np.abs(div).max() <= 2e-3 * grad_scale  # float32 output, central differences
    rms = np.sqrt((u * u).sum(0).mean())
    assert abs(rms - 0.5) < 1e-3
    # seeded: the same seed gives the same field
    again = bench.curl_noise_velocity(allp[0], allp[1], allp[2], 20260115, 0.5, 5.1)
    assert np.array_equal(again, u.astype(np.float32))


def test_kelvin_map_anchors():
    assert abs(bench.kelvin_to_model(273.15) - 0.105) < 1e-12
    assert abs(bench.kelvin_to_model(333.15) - 1.105) < 1e-12
    assert bench.kelvin_to_model(270.0) > 0 and bench.kelvin_to_model(300.0) < 1.15
