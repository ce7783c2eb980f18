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


def oracle_step_frame(st, dt_frame, clock, sun_activity, currents, substeps=1):
    """FluidSolver._stepSync (js/solver.js:2670-2700) over the oracle's stage
    functions: the 25 Hz heat / evaporation / currents tick (:2672-2681), the
    CFL substep loop (:2682-2694) and the frame-end passes in their order
    (:2696-2699). `clock` holds the host-side simTime and heatAcc, as the
    reference keeps them on the solver object. Returns the substep lengths."""
    import ctypes
    clock["simTime"] += dt_frame
    st.set_params(simTime=clock["simTime"])
    clock["heatAcc"] += dt_frame
    if clock["heatAcc"] >= 1.0 / 25:
        tick = min(clock["heatAcc"], 1.0 / 12)
        st.call("oc_update_heat", ctypes.c_double(tick))
        if sun_activity > 0:
            st.call("oc_update_evaporation", ctypes.c_double(tick))
        if currents > 0:
            st.call("oc_update_currents", ctypes.c_double(tick))
        clock["heatAcc"] = 0.0
        clock["ticks"] = clock.get("ticks", 0) + 1
    remaining, n, base = dt_frame, 0, dt_frame / max(1, substeps)
    dts = []
    while remaining > 1e-6 and n < 10:
        # oc_dt_law is min(cfl dx / max(umax, 0.08), 1/100) floored at 1e-5
        dt = max(min(base, st.call("oc_dt_law")), 1e-5)
        st.call("oc_substep", ctypes.c_double(dt))
        dts.append(dt)
        remaining -= dt
        n += 1
    st.call("oc_vapor_collisions")
    st.call("oc_update_phase", ctypes.c_double(dt_frame))
    st.call("oc_ice_contacts")
    st.call("oc_push_surface")
    return dts
