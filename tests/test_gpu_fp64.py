"""Error of each GPU stage against an FP64 run of the same stage.

The oracle built with double storage (oracle/_build/liboracle64.so) runs
every substep stage from the reference's dumped input of that stage. The
GPU runs the same stage from the same input. Two errors are reported per
output field, both relative to the field's largest magnitude:
  - GPU - FP64;
  - reference - FP64, the reference's own float-storage error.
The bar is the stage bar of the golden comparison (2e-5 of the largest
magnitude; pressure 2e-5 after the same red-black sweeps). Particle fields are
compared only where both runs took the same discrete branch (flag, cooldown),
as in the golden comparison. The numbers land in the test log.
"""
import ctypes

import numpy as np
import pytest

import golden
from gpu_scenes import gpu_from_golden
from scenes import oracle_from_golden, PART_KEYS

pytestmark = pytest.mark.gpu

SCENES = ["cfg1_32", "micro_16", "thermal_32"]
TOL = 2e-5


def _rel(a, b, scale):
    return float(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).max()) / scale


def _report(stage, k, gpu, ref, f64, mask=None):
    if mask is not None:
        gpu, ref, f64 = gpu[mask], ref[mask], f64[mask]
    scale = max(float(np.abs(f64).max()), 1e-30)
    eg, er = _rel(gpu, f64, scale), _rel(ref, f64, scale)
    print("  %-10s %-6s gpu-fp64 %.3e   reference-fp64 %.3e" % (stage, k, eg, er))
    assert eg <= TOL, (stage, k, eg)
    return eg


def _grid_in(s, o, g, tag, names=("u", "v", "w")):
    for k in names:
        s.upload(k, g[tag + k])
        o.arrays[k][:] = g[tag + k]


@pytest.mark.parametrize("scene", SCENES)
def test_stage_errors_vs_fp64(scene):
    g = golden.load(scene)
    s = gpu_from_golden(g)
    o = oracle_from_golden(g, bits=64)
    dt = g.meta["stage_dt"]
    n = len(g["in_px"])
    print("\n[%s]" % scene)
    # mark + P2G from the reference's input particles
    s.stage("clear")
    s.stage("rasterize")
    s.stage("p2g")
    o.call("oc_clear_faces")
    o.call("oc_rasterize_static")
    o.call("oc_mark_fluid", 0, n)
    assert np.array_equal(o["cellType"], g["mark_cellType"])
    o.call("oc_p2g")
    for k in ("u", "v", "w", "uW", "vW", "wW"):
        _report("p2g", k, s.download(k), g["p2g_" + k], o[k])
    # boundary conditions
    _grid_in(s, o, g, "p2g_")
    s.stage("bc")
    o.call("oc_apply_bc")
    for k in "uvw":
        _report("bc", k, s.download(k), g["bc_" + k], o[k])
    # viscosity
    _grid_in(s, o, g, "bc_")
    s.stage("viscosity", dt)
    o.call("oc_viscosity", ctypes.c_double(dt))
    for k in "uvw":
        _report("viscosity", k, s.download(k), g["visc_" + k], o[k])
    # gravity (+ BC)
    _grid_in(s, o, g, "visc_")
    s.set(gRamp=g.meta["gRamp_before_gravity"])
    o.set_params(gRamp=g.meta["gRamp_before_gravity"])
    s.stage("gravity", dt)
    o.call("oc_apply_gravity", ctypes.c_double(dt))
    o.call("oc_apply_bc")
    for k in "uvw":
        _report("gravity", k, s.download(k), g["grav_" + k], o[k])
    # pressure: 40 red-black sweeps on both from the reference's start q
    _grid_in(s, o, g, "grav_")
    s.upload("q", g["in_q"])
    o.arrays["q"][:] = g["in_q"]
    s.set(pressureIters=40)
    s.stage("pressure", dt)
    o.set_params(pressureIters=40, sorRedBlack=1)
    o.call("oc_pressure_solve")
    q64 = o["q"].copy()
    # the reference's q is lexicographic SOR: shown for its size, not gated
    scale = float(np.abs(q64).max())
    print("  %-10s %-6s gpu-fp64 %.3e   (red-black on both sides)" % ("pressure", "q", _rel(s.download("q"), q64, scale)))
    assert _rel(s.download("q"), q64, scale) <= TOL
    # projection with the reference's pressure
    _grid_in(s, o, g, "grav_")
    s.upload("q", g["pres_q"])
    o.arrays["q"][:] = g["pres_q"]
    o.arrays["cellPhi"][:] = g["pres_q"]
    s.stage("project")
    o.call("oc_project")
    for k in "uvw":
        _report("project", k, s.download(k), g["proj_" + k], o[k])
    # vorticity confinement
    if g.meta["stage_params"]["vorticity"] > 0:
        _grid_in(s, o, g, "proj_")
        s.stage("vorticity", dt)
        o.call("oc_vorticity", ctypes.c_double(dt))
        for k in "uvw":
            _report("vorticity", k, s.download(k), g["vort_" + k], o[k])
    # extrapolation from the reference's clamped field and P2G weights
    _grid_in(s, o, g, "clamp_")
    _grid_in(s, o, g, "p2g_", ("uW", "vW", "wW"))
    s.stage("extrapolate")
    o.call("oc_extrapolate")
    for k in "uvw":
        _report("extrap", k, s.download(k), g["ext_" + k], o[k])
    # G2P on the reference's grids
    _grid_in(s, o, g, "ext_")
    for k in "uvw":
        s.upload(k + "O", g["bc_" + k])
        o.arrays[k + "O"][:] = g["bc_" + k]
    for k in PART_KEYS:
        s.upload(k, g["in_" + k])
        o.arrays[k][:n] = g["in_" + k]
    s.upload("pflag", g["mark_pflag"])
    o.arrays["pflag"][:n] = g["mark_pflag"]
    s.stage("g2p", dt)
    o.call("oc_g2p", ctypes.c_double(dt))
    same = (s.download("pflag") == o["pflag"][:n]) & (s.download("pcool") == o["pcool"][:n])
    same &= g["g2p_pflag"] == o["pflag"][:n]
    for k in ("pvx", "pvy", "pvz"):
        _report("g2p", k, s.download(k), g["g2p_" + k], o[k][:n], same)
    # advection from the reference's G2P state
    for k in PART_KEYS:
        s.upload(k, g["g2p_" + k])
        o.arrays[k][:n] = g["g2p_" + k]
    for k in ("advU", "advV", "advW"):
        s.upload(k, g["g2p_" + k].astype(np.float32))
        o.arrays[k][:n] = g["g2p_" + k]
    if "rng_before" in g.meta:
        rb = g.meta["rng_before"]
        s.set_rng(rb["cfS"], rb["bfS"], rb["evS"])
        o.set_params(cfS=rb["cfS"], bfS=rb["bfS"], evS=rb["evS"])
    s.stage("advect", dt)
    o.call("oc_advect", ctypes.c_double(dt))
    same = (s.download("pflag") == o["pflag"][:n]) & (g["adv_pflag"] == o["pflag"][:n])
    same &= s.download("pcool") == o["pcool"][:n]
    for k in ("px", "py", "pz", "pvx", "pvy", "pvz"):
        _report("advect", k, s.download(k), g["adv_" + k], o[k][:n], same)
    print("  particles on the same branch in all three runs: %d / %d" % (int(same.sum()), n))
    s.close()
