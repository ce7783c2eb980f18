"""Pins the oracle's frame tick against the reference on the thermal scene
(tests/golden/gen_golden.js, scene C): _updateHeat, _updateEvaporation,
_updateCurrents and _updatePhaseChanges, each from the reference's input
state. Flags, counts and stream states are exact; the stages that call
exp/log/cos carry the libm-vs-V8 last-ulp difference, so float fields are
compared to 1e-6."""
import numpy as np

import golden
from scenes import oracle_from_golden, PART_KEYS, EXTRA_KEYS, tick_params


def load_tick(tag):
    g = golden.load("thermal_32")
    st = oracle_from_golden(g, tag=tag)
    st.arrays["cellType"][:] = g["t0_cellType"]
    st.arrays["dens"][:] = g["t0_dens"]
    for k in "uvw":
        st.arrays[k][:] = g["t0_" + k]
    st.set_params(**tick_params(g))
    st.set_params(evS=g.meta["tick"]["evS_before"])
    return g, st


def close(st, g, tag, keys, n, tol=1e-6):
    for k in keys:
        a, b = st[k][:n].astype(np.float64), g[tag + k].astype(np.float64)
        assert np.abs(a - b).max() <= tol, (tag + k, np.abs(a - b).max())


def test_heat_tick():
    g, st = load_tick("t0_")
    n = len(g["t0_px"])
    st.call("oc_update_heat", g.meta["tick"]["dtTick"])
    assert np.array_equal(st["pDepth"][:n], g["t1_pDepth"])
    close(st, g, "t1_", ["pT", "pLight", "pvx", "pvy", "pvz"], n)


def test_heat_conserves_conduction_energy():
    """optimization.test.js:3-8: conduction only moves heat between particles:
    the same tick with heatK 1 and heatK 0 leaves the same total."""
    g = golden.load("thermal_32")
    n = len(g["t0_px"])
    tot = []
    for k in (0.0, 1.0):
        _, st = load_tick("t0_")
        st.set_params(heatK=k)
        st.call("oc_update_heat", 0.04)
        tot.append(float(st["pT"][:n].astype(np.float64).sum()))
    assert abs(tot[0] - tot[1]) < 1e-5 * max(abs(tot[0]), 1.0)


def test_evaporation_stream_exact():
    g, st = load_tick("t1_")
    n = len(g["t1_px"])
    st.set_params(evS=g.meta["tick"]["evS_before"])
    st.call("oc_update_evaporation", g.meta["tick"]["dtTick"])
    assert st.get_i("evapCount") == g.meta["tick"]["evapCount"]
    assert st.get_i("evS") == g.meta["tick"]["evS_after"]
    assert np.array_equal(st["pflag"][:n], g["t2_pflag"])
    close(st, g, "t2_", ["pvx", "pvy", "pvz", "pAir"], n)


def test_currents():
    g, st = load_tick("t2_")
    n = len(g["t2_px"])
    st.call("oc_update_currents", g.meta["tick"]["dtTick"])
    close(st, g, "t3_", ["pvx", "pvy", "pvz"], n)


def test_phase_changes():
    g, st = load_tick("t3_")
    n = len(g["t3_px"])
    st.call("oc_update_phase", 1 / 60)
    assert np.array_equal(st["pflag"][:n], g["t4_pflag"])
    close(st, g, "t4_", ["pvx", "pvy", "pvz", "pAir", "pWx", "pWy", "pWz", "pPh"], n)
