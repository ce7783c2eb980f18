"""Loads a golden scene into a GPU solver context (HIP library through its C
ABI), mirroring scenes.oracle_from_golden for the CPU oracle."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "watersim-v1_amd"))

import watersim_hip  # noqa: E402

from scenes import PART_KEYS, EXTRA_KEYS, tick_params  # noqa: E402


def gpu_from_golden(g, tag="in_", capacity=None):
    w = g.meta["world"]
    n = len(g[tag + "px"])
    s = watersim_hip.Solver(w["nx"], w["ny"], w["nz"], w["dx"], capacity or n)
    s.set(coreR=w["coreR"], oceanR=w["oceanR"], spacing=w["spacing"], particleVolume=w["particleVolume"],
          atmosphereH=w["atmosphereH"], dv=w["dv"], tn=w["tn"])
    sp = g.meta["stage_params"]
    s.set(pic=sp["pic"], omega=sp["omega"], pressureIters=sp["pressureIters"], gravity=sp["gravity"],
          maxSpeed=sp["maxSpeed"], viscosity=sp["viscosity"], vorticity=sp["vorticity"], cfl=sp["cfl"])
    s.upload("staticTypes", g["static_types"])
    s.upload("gu", g["gu"])
    s.upload("gv", g["gv"])
    s.upload("gw", g["gw"])
    s.upload("terrainR", g["terrain_R"])
    s.set(nP=n)
    for k in PART_KEYS + EXTRA_KEYS:
        if tag + k in g:
            s.upload(k, g[tag + k])
    if tag == "in_":
        s.upload("q", g["in_q"])
        s.set(gRamp=sp["gRamp"])
    tp = tick_params(g)
    if tp:
        s.upload("terrainSolid", g["terrain_solid"])
        s.set(**tp)
        rb = g.meta["rng_before"]
        s.set_rng(rb["cfS"], rb["bfS"], rb["evS"])
        if any(g[tag + "pflag"] >= 2) if tag + "pflag" in g else False:
            s.set(hasVapor=1)
    return s


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    return np.abs(a - b).max() / scale


ORACLE_PARTS = ["px", "py", "pz", "pvx", "pvy", "pvz", "pT", "pAir", "pflag", "pcool", "pLight", "pDepth", "pWx",
                "pWy", "pWz", "pPh"]


def oracle_from_solver(s, world, bits=32):
    """The CPU oracle on a GPU context's current state: the world the context
    was built from, every particle field, q, the gravity ramp, umax and the
    advection streams, and the context's solver settings."""
    import cpu_oracle as oracle
    n = world["nx"]
    st = oracle.OracleState({"nx": n, "ny": n, "nz": n, "dx": world["dx"], "coreR": world["coreR"],
                             "oceanR": world["oceanR"], "spacing": world["spacing"],
                             "particleVolume": world["particleVolume"], "dv": world["dv"], "tn": world["tn"],
                             "atmosphereH": s.get("atmosphereH")},
                            bits=bits)
    st.set_static(world["solid"], world["R"], world["staticTypes"], world["gu"], world["gv"], world["gw"])
    nP = int(s.get("nP"))
    st.set_particles(nP, **{k: s.download(k) for k in ORACLE_PARTS})
    st.arrays["q"][:] = s.download("q")
    st.set_params(**{k: s.get(k) for k in ("pic", "omega", "gravity", "maxSpeed", "vorticity", "cfl", "viscosity",
                                             "gRamp", "umax")})
    cf, bf, ev = s.get_rng()
    st.set_params(cfS=cf, bfS=bf, evS=ev, Rlo=world["Rlo"], Rhi=world["Rhi"])
    return st
