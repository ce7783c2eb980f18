"""Balls on the GPU against the reference's ball scene (32^3 planet, five
balls) and the CPU oracle.

The ball kernels evaluate the reference's double formulas in FP64 without
fused operations, so the stages that see the same inputs are bit-exact:
the moving-solid raster (cell types, solid velocities), the particle
push-out and the integrator (waterline histogram, waterline, all nine
numbers of every ball). The full substep carries the FP32 grid differences
of the other stages into the balls through the drag term; its bar is the
substep bar of the other scenes (particles) and 1e-4 relative on the balls.
"""
import numpy as np
import pytest

import golden
from gpu_scenes import gpu_from_golden
from scenes import oracle_from_golden

pytestmark = pytest.mark.gpu

PART = ["px", "py", "pz", "pvx", "pvy", "pvz"]


def ball_scene(tag="in_", balls="in_balls"):
    g = golden.load("balls_32")
    s = gpu_from_golden(g, tag=tag)
    s.upload("terrainSolid", g["terrain_solid"])
    s.set_balls(g[balls])
    return g, s


def report_cells(got, want, g):
    d = np.nonzero(got != want)[0]
    ball = (g["rast_cellType"] == 2) & (g["static_types"] != 2)
    for c in d[:8]:
        print("   cell %d: got %d want %d static %d ball-cell %d" % (c, got[c], want[c], g["static_types"][c], ball[c]))
    return len(d)


def test_ball_raster_bit_exact():
    g, s = ball_scene()
    s.stage("clear")
    s.stage("rasterize")
    got = s.download("cellType")
    assert report_cells(got, g["mark_cellType"], g) == 0
    for k in ["svx", "svy", "svz"]:
        assert np.array_equal(s.download(k), g["rast_" + k]), k
    print("  ball cells %d, moving-solid velocities bit-exact"
          % int(((g["rast_cellType"] == 2) & (g["static_types"] != 2)).sum()))
    s.close()


def test_ball_push_out_bit_exact():
    g, s = ball_scene()
    s.stage("clear")
    s.stage("rasterize")
    s.stage("push_balls")
    moved = 0
    for k in PART:
        got = s.download(k)
        assert np.array_equal(got, g["pushb_" + k]), k
        moved += int((got != g["in_" + k]).sum())
    print("  pushed particle components %d, bit-exact" % moved)
    s.close()


def test_ball_integrator_bit_exact():
    """_updateBalls on the reference's own post-advection particles and
    extrapolated grid."""
    g, s = ball_scene(tag="adv_", balls="in_balls")
    for k in "uvw":
        s.upload(k, g["ext_" + k])
    s.stage("update_balls", g.meta["stage_dt"])
    hist = s.download("ballHist").reshape(-1, 25)
    assert np.array_equal(hist[:, :24].ravel().astype(np.float32), g["ball_hist"])
    assert np.array_equal(s.download("ballSurf"), g["ball_surf"])
    got = s.balls()
    want = g["out_balls"].reshape(-1, 9)
    print("  balls max |diff| %.3g" % np.abs(got - want).max())
    assert np.array_equal(got, want)
    s.close()


def test_substep_with_balls_against_oracle():
    g, s = ball_scene()
    st = oracle_from_golden(g)
    st.set_balls(g["in_balls"])
    dt = g.meta["stage_dt"]
    s.substep(dt)
    st.call("oc_substep", dt)
    n = len(g["in_px"])
    for k in PART:
        ref = st[k][:n].astype(np.float64)
        err = np.abs(s.download(k) - ref)
        bad = int((err > 1e-4 * max(1.0, np.abs(ref).max())).sum())
        print("  %-4s max |err| %.3g, beyond 1e-4: %d" % (k, err.max(), bad))
        assert bad <= max(1, n // 1000), k
    got, want = s.balls(), st.balls()
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1.0)
    print("  balls max rel err %.3g" % rel.max())
    assert rel.max() <= 1e-4
    s.close()


def test_ball_trajectory_follows_reference():
    """20 substeps from the fill state with the dt law on the host, ball
    states against the reference's recorded trajectory."""
    g, s = ball_scene(tag="fill_", balls="fill_balls")
    s.upload("q", g["fill_q"])
    s.set(gRamp=g.meta["fill"]["gRamp"])
    rec = g.meta["trajectory"]
    tb = g["traj_balls"].reshape(len(rec), -1, 9)
    worst = 0.0
    for i in range(len(rec)):
        s.substep(rec[i]["dt"])
        got = s.balls()
        pos = np.abs(got[:, :3] - tb[i][:, :3]).max()
        vel = np.abs(got[:, 3:6] - tb[i][:, 3:6]).max()
        worst = max(worst, pos)
        assert pos <= 2e-3 and vel <= 5e-2, (i, pos, vel)
    print("  20 substeps: worst ball position difference %.3g" % worst)
    s.close()
