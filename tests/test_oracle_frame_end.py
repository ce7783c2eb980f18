"""The frame-end passes of the reference (js/solver.js:3290-3587) in the
oracle, bit-exact against the reference's own run on the thermal scene,
whose state was perturbed so every branch runs (particles pushed into the
rock, over the atmosphere cap and out of the domain)."""
import numpy as np

import golden
from scenes import oracle_from_golden, PART_KEYS, EXTRA_KEYS


def load(tag):
    g = golden.load("thermal_32")
    st = oracle_from_golden(g, tag=tag)
    return g, st


def test_push_surface_bit_exact():
    g, st = load("fe2_")
    n = len(g["fe2_px"])
    st.call("oc_push_surface")
    changed = 0
    for k in PART_KEYS + EXTRA_KEYS:
        if "fe3_" + k not in g:
            continue
        got, want = st[k][:n], g["fe3_" + k]
        assert np.array_equal(got, want), (k, int(np.sum(got != want)))
        changed += int(np.sum(g["fe2_" + k] != want))
    assert changed > 100  # the pass did real work on this state


def test_vapor_collisions_bit_exact():
    g, st = load("fe0_")
    n = len(g["fe0_px"])
    fe = g.meta["frame_end"]
    sp = fe["sunPos"]
    st.set_params(spinOn=1, hasSun=1, sunX=sp[0], sunY=sp[1], sunZ=sp[2])
    st.call("oc_vapor_collisions")
    assert st.get_i("vaporCount") == fe["vaporCount"]
    assert st.get_i("nightVaporCount") == fe["nightVaporCount"]
    changed = 0
    for k in PART_KEYS + EXTRA_KEYS:
        if "fe1_" + k not in g:
            continue
        got, want = st[k][:n], g["fe1_" + k]
        assert np.array_equal(got, want), (k, int(np.sum(got != want)))
        changed += int(np.sum(g["fe0_" + k] != want))
    assert changed > 100


def test_ice_contacts_bit_exact():
    g, st = load("fe1_")
    n = len(g["fe1_px"])
    st.call("oc_ice_contacts")
    changed = 0
    for k in PART_KEYS + EXTRA_KEYS:
        if "fe2_" + k not in g:
            continue
        got, want = st[k][:n], g["fe2_" + k]
        assert np.array_equal(got, want), (k, int(np.sum(got != want)))
        changed += int(np.sum(g["fe1_" + k] != want))
    assert changed > 20
