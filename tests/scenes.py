"""Scene loading shared by the oracle and GPU parity tests: builds an oracle
state from a golden fixture's world and the particle state entering the
dumped substep."""
import numpy as np

import golden
import cpu_oracle as oracle

PART_KEYS = ["px", "py", "pz", "pvx", "pvy", "pvz", "pT", "pAir", "pflag", "pcool"]
EXTRA_KEYS = ["pLight", "pDepth", "pWx", "pWy", "pWz", "pPh"]


def tick_params(g):
    """Frame-tick settings of a scene (sun, thermal, evaporation, currents)."""
    t = g.meta.get("tick")
    if not t:
        return {}
    sp = t["sunPos"]
    return dict(hasSun=1, sunX=sp[0], sunY=sp[1], sunZ=sp[2], sunActivity=t["sunActivity"],
                evapIntensity=t["evapIntensity"], heatK=t["heatK"], Tamb=t["Tamb"], shadeCool=t["shadeCool"],
                snowT=t["snowT"], meltT=t["meltT"], rainT=t["rainT"], cloudT=t["cloudT"], cloudP=t["cloudP"],
                rainLiftAge=t["rainLiftAge"], currents=t["currents"], Rlo=t["Rlo"], Rhi=t["Rhi"],
                simTime=t["simTime"])


def world_of(g):
    w = dict(g.meta["world"])
    return w


def particles(g, tag):
    return {k: g[tag + k] for k in PART_KEYS if tag + k in g}


def oracle_from_golden(g, tag="in_", bits=32):
    w = world_of(g)
    st = oracle.OracleState(w, bits=bits)
    st.set_static(g["terrain_solid"], g["terrain_R"], g["static_types"], g["gu"], g["gv"], g["gw"])
    parts = particles(g, tag)
    for k in EXTRA_KEYS:
        if tag + k in g:
            parts[k] = g[tag + k]
    n = len(parts["px"])
    st.set_particles(n, **parts)
    tp = tick_params(g)
    if tp:
        st.set_params(**tp)
        rb = g.meta["rng_before"]
        st.set_params(cfS=rb["cfS"], bfS=rb["bfS"], evS=rb["evS"])
    sp = g.meta["stage_params"]
    st.set_params(pic=sp["pic"], omega=sp["omega"], pressureIters=sp["pressureIters"],
                  gravity=sp["gravity"], maxSpeed=sp["maxSpeed"], viscosity=sp["viscosity"],
                  vorticity=sp["vorticity"], cfl=sp["cfl"], gRamp=sp["gRamp"])
    if tag == "in_":
        st.arrays["q"][:] = g["in_q"]
    return st
