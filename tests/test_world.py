"""The host world builder (ws_world_build) against the reference's worlds in
the golden fixtures: same seed, same terrain voxels, radii, static raster,
gravity faces and particle fill, bit for bit. Needs no GPU."""
import os
import sys

import numpy as np
import pytest

import golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "watersim-v1_amd"))
import watersim_hip  # noqa: E402

SCENES = [("cfg1_32", 20260114, 10000), ("micro_16", 20260115, 2000)]


@pytest.mark.parametrize("scene,seed,target", SCENES)
def test_world_matches_reference(scene, seed, target):
    g = golden.load(scene)
    wm = g.meta["world"]
    w = watersim_hip.build_world(wm["nx"], wm["dx"], wm["coreR"], wm["oceanR"] - wm["coreR"], seed, target)
    assert w["tn"] == wm["tn"]
    assert np.array_equal(w["solid"], g["terrain_solid"])
    assert np.array_equal(w["R"], g["terrain_R"])
    assert np.array_equal(w["staticTypes"], g["static_types"])
    for k in ("gu", "gv", "gw"):
        assert np.array_equal(w[k], g[k]), k
    assert w["nP"] == len(g["fill_px"])
    for k in ("px", "py", "pz"):
        assert np.array_equal(w[k], g["fill_" + k]), k
    assert abs(w["spacing"] - wm["spacing"]) <= 1e-15 * wm["spacing"]
    assert w["landFrac"] == wm["landFrac"]


@pytest.mark.parametrize("scene", ["cfg1_32", "micro_16", "thermal_32"])
def test_planet_gravity_matches_reference(scene):
    """ws_planet_gravity (what the facade's _buildPlanetGravity calls) equals
    the reference's _buildPlanetGravity faces (solver.js:1177-1210) for the
    scene's ocean radius, bit for bit."""
    g = golden.load(scene)
    wm = g.meta["world"]
    gu, gv, gw = watersim_hip.planet_gravity(wm["nx"], wm["dx"], wm["oceanR"])
    assert np.array_equal(gu, g["gu"]) and np.array_equal(gv, g["gv"]) and np.array_equal(gw, g["gw"])
