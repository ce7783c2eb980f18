"""Scene loading shared by the oracle and GPU parity tests: builds an oracle
state from a golden fixture's world and the particle state entering the
dumped substep."""
import numpy as np

import golden
import cpu_oracle as oracle

PART_KEYS = ["px", "py", "pz", "pvx", "pvy", "pvz", "pT", "pAir", "pflag", "pcool"]


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
    n = len(parts["px"])
    st.set_particles(n, **parts)
    sp = g.meta["stage_params"]
    st.set_params(pic=sp["pic"], omega=sp["omega"], pressureIters=sp["pressureIters"],
                  gravity=sp["gravity"], maxSpeed=sp["maxSpeed"], viscosity=sp["viscosity"],
                  vorticity=sp["vorticity"], cfl=sp["cfl"], gRamp=sp["gRamp"])
    if tag == "in_":
        st.arrays["q"][:] = g["in_q"]
    return st
