"""The oracle's metaball field (Metaballs.build, js/surface.js:178-284)
against the reference: the field contoured by the oracle's marching
tetrahedra gives the reference's mesh bit for bit (vertex count, FP64 volume,
positions, normals) on a settled 32^3 planet over a sweep of tension and
detail settings, and on the 7^3 particle block of test/metaball.test.js.
The lattice / radius logic is pinned the same way (a wrong lattice size or
radius changes the mesh). The contour level is 1.5 x^1.5, and x^1.5 comes
from the platform's pow: the C library's and V8's Math.pow may differ in
the last bit, so the meshes are built at the reference's level and the
oracle's own level is held to 2 ulp of it."""
import numpy as np
import pytest

import golden
import cpu_oracle as oracle


def mesh(px, py, pz, fl, nx, dx, spacing, tension, detail, iso_ref=None):
    m, mdx, R, rad, iso = oracle.metaball_params(nx, dx, spacing, tension,
                                                 float("nan") if detail is None else detail)
    field = oracle.metaball_field(px, py, pz, fl, m, mdx, R, rad)
    return oracle.mt_build(field, m, mdx, iso if iso_ref is None else iso_ref), iso


def test_planet_meshes_bit_exact():
    g = golden.load("metaballs")
    P = g.meta["planet"]
    px, py, pz, fl = g["planet_px"], g["planet_py"], g["planet_pz"], g["planet_pflag"]
    for ci, case in enumerate(g.meta["planet_cases"]):
        (cnt, vol, pos, nrm), iso = mesh(px, py, pz, fl, P["nx"], P["dx"], P["spacing"], case["tension"],
                                         case["detail"], case["iso"])
        assert abs(iso - case["iso"]) <= 2 * np.spacing(case["iso"]), ci
        assert cnt == case["count"], ci
        assert vol == case["volume"], ci
        if "planet_pos_%d" % ci in g:
            assert np.array_equal(pos, g["planet_pos_%d" % ci]), ci
            assert np.array_equal(nrm, g["planet_nrm_%d" % ci]), ci


def test_reference_metaball_block():
    g = golden.load("metaballs")
    S = g.meta["stub"]
    n = S["nP"]
    fl = np.zeros(n, np.uint8)
    isos = {c["tension"]: c["iso"] for c in g.meta["planet_cases"]}
    for case in S["cases"]:
        (cnt, vol, _, _), _ = mesh(g["stub_px"], g["stub_py"], g["stub_pz"], fl, S["nx"], S["dx"], S["spacing"],
                                   case["tension"], case["detail"], isos.get(case["tension"]))
        assert cnt == case["count"] and vol == case["volume"], case


@pytest.mark.parametrize("tension,detail,m", [(0.5, None, 32), (0.5, 0.6, 19), (0.5, 0.1, 16), (0.5, 9.0, 51),
                                              (0.5, -1.0, 32), (0.5, float("inf"), 32)])
def test_lattice_of_detail(tension, detail, m):
    """detail clamps to 0.5..1.6 of the solver lattice; missing, negative or
    non-finite detail is the solver lattice itself (mdx == dx exactly)."""
    got = oracle.metaball_params(32, 0.159375, 0.0975, tension, float("nan") if detail is None else detail)
    assert got[0] == m
    if m == 32:
        assert got[1] == 0.159375


def test_iso_and_radius_of_tension():
    """iso = 1.5 ((1.6 + 2.2 t) / 1.6)^1.5 rises with tension and clamps
    outside [0, 1]; the kernel radius is (1.6 + 2.2 t) spacings."""
    isos = [oracle.metaball_params(16, 1 / 16, 1 / 64, t)[4] for t in (-1, 0, 0.5, 1, 2)]
    assert isos[0] == isos[1] == 1.5 and isos[3] == isos[4]
    assert isos[1] < isos[2] < isos[3]
    assert abs(isos[3] - 1.5 * (3.8 / 1.6) ** 1.5) < 1e-12
    R = oracle.metaball_params(16, 1 / 16, 1 / 64, 1.0)[2]
    assert abs(R - 3.8 / 4) < 1e-12
