"""Metaballs.build (js/surface.js:178-284) on the GPU against the oracle and
the reference's meshes, plus the properties the reference's own test
(test/metaball.test.js) asserts.

Bar: the lattice size and contour level equal the reference's (level within
2 ulp: the platform pow); the field equals the oracle's to 1e-5 of its
largest value (the GPU sums exactly in fixed point, the reference rounds to
float after each of up to a few hundred additions, in particle order); vertex counts within 0.2 % and volumes within 1e-5
relative of the reference; on the GPU's own field the mesh equals the
oracle mesher's bit for bit (the volume to 1e-9: another summation order).
"""
import numpy as np
import pytest

import golden
import cpu_oracle as oracle
from gpu_scenes import watersim_hip

pytestmark = pytest.mark.gpu


def context(n, dx, spacing, px, py, pz, fl):
    s = watersim_hip.Solver(n, n, n, dx, max(len(px), 1))
    s.set(spacing=spacing, nP=len(px))
    if len(px):
        s.upload("px", px)
        s.upload("py", py)
        s.upload("pz", pz)
        s.upload("pflag", fl)
    return s


def test_planet_metaballs_against_reference():
    g = golden.load("metaballs")
    P = g.meta["planet"]
    px, py, pz, fl = g["planet_px"], g["planet_py"], g["planet_pz"], g["planet_pflag"]
    s = context(P["nx"], P["dx"], P["spacing"], px, py, pz, fl)
    for ci, case in enumerate(g.meta["planet_cases"]):
        det = float("nan") if case["detail"] is None else case["detail"]
        cnt, vol, pos, nrm, iso, m = s.metaballs_build(case["tension"], det)
        om, omdx, oR, orad, oiso = oracle.metaball_params(P["nx"], P["dx"], P["spacing"], case["tension"], det)
        assert m == om
        assert abs(iso - case["iso"]) <= 2 * np.spacing(case["iso"])
        field = s.download("metaballField")
        ref = oracle.metaball_field(px, py, pz, fl, om, omdx, oR, orad)
        ferr = np.abs(field.astype(np.float64) - ref).max()
        assert ferr <= 1e-5 * ref.max(), (ci, ferr)
        dc = abs(cnt - case["count"]) / case["count"]
        dv = abs(vol - case["volume"]) / case["volume"]
        msg = "  t=%.2f d=%s m=%d: count %d (ref %d), volume rel %.2e, field max |err| %.2e" % (
            case["tension"], case["detail"], m, cnt, case["count"], dv, ferr)
        # the mesher on the GPU's own field against the oracle's mesher on
        # the same field (the field itself differs from the reference's by
        # float roundings, which move vertices on nearly flat edges)
        ocnt, ovol, opos, onrm = oracle.mt_build(field, om, omdx, iso)
        same = ocnt == cnt and np.array_equal(pos, opos) and np.array_equal(nrm, onrm)
        msg += ", same-field mesh: count %d vs %d, bit-exact %s" % (cnt, ocnt, same)
        print(msg)
        assert same and abs(ovol - vol) <= 1e-9 * ovol, ci
        assert dc <= 2e-3 and dv <= 1e-5, ci
    s.close()


def test_reference_block_properties():
    """The reference test's 7^3 block: a closed mesh with unit normals that
    point outward, deterministic, volume growing with tension, detail
    omitted / NaN / non-positive equal to detail 1, absurd details clamped,
    and the counts and volumes of the reference."""
    g = golden.load("metaballs")
    S = g.meta["stub"]
    n = S["nP"]
    fl = np.zeros(n, np.uint8)
    s = context(S["nx"], S["dx"], S["spacing"], g["stub_px"], g["stub_py"], g["stub_pz"], fl)
    prev = -1.0
    for case in S["cases"]:
        det = float("nan") if case["detail"] is None else case["detail"]
        cnt, vol, pos, nrm, iso, m = s.metaballs_build(case["tension"], det)
        assert cnt == case["count"], case
        assert abs(vol - case["volume"]) <= 1e-5 * case["volume"], case
        if case["detail"] is None:
            assert vol > prev
            prev = vol
    cnt, vol, pos, nrm, _, _ = s.metaballs_build(0.0)
    assert cnt > 0 and cnt % 3 == 0
    p, q = pos.reshape(-1, 3), nrm.reshape(-1, 3)
    assert np.all(np.isfinite(p)) and np.all(np.isfinite(q))
    assert np.abs(np.linalg.norm(q, axis=1) - 1).max() < 1e-5
    assert np.all(((p - 0.5) * q).sum(1) > -0.5)
    again = s.metaballs_build(0.0)
    assert again[0] == cnt and np.array_equal(again[2], pos)
    base = s.metaballs_build(0.5)
    for d in (-1.0, 0.0, float("nan"), float("inf")):
        other = s.metaballs_build(0.5, d)
        assert other[0] == base[0] and other[1] == base[1], d
    assert s.metaballs_build(0.5, 0.1)[5] == s.metaballs_build(0.5, 0.5)[5] == 8
    assert s.metaballs_build(0.5, 9.0)[5] == s.metaballs_build(0.5, 1.6)[5] == 26
    s.close()


def test_only_liquid_particles_count_and_empty_gives_nothing():
    g = golden.load("metaballs")
    S = g.meta["stub"]
    n = S["nP"]
    px, py, pz = g["stub_px"], g["stub_py"], g["stub_pz"]
    fl = np.zeros(n, np.uint8)
    fl[::3] = 2  # vapor: not part of the liquid skin
    s = context(S["nx"], S["dx"], S["spacing"], px, py, pz, fl)
    s.metaballs_build(0.5)
    f1 = s.download("metaballField")
    s.close()
    keep = fl == 0
    s = context(S["nx"], S["dx"], S["spacing"], px[keep], py[keep], pz[keep], fl[keep])
    s.metaballs_build(0.5)
    assert np.array_equal(f1, s.download("metaballField"))
    s.close()
    s = context(S["nx"], S["dx"], S["spacing"], px[:0], py[:0], pz[:0], fl[:0])
    cnt, vol, _, _, _, _ = s.metaballs_build(0.5)
    assert cnt == 0 and vol == 0
    s.close()
