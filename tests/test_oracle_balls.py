"""The CPU oracle's balls (moving solids) against the reference run on the
ball scene of tests/golden (32^3 planet, five balls: floating, dropped, an
overlapping pair, one sunk into a continent).

Pins, all bit-exact: the moving-solid raster (cell types and solid
velocities, solver.js:918-938), the particle push-out (solver.js:1005-1026),
the full substep with the ball integrator (solver.js:2344-2564: the
waterline histogram, the surface level, every ball's state) and a 20-substep
replay of the ball trajectories from the fill state."""
import numpy as np

import golden
from scenes import oracle_from_golden, PART_KEYS


def state_with_balls(g, tag="in_"):
    st = oracle_from_golden(g, tag=tag)
    st.set_balls(g[tag + "balls"])
    return st


def test_ball_raster_bit_exact():
    g = golden.load("balls_32")
    st = state_with_balls(g)
    st.call("oc_clear_faces")
    st.call("oc_rasterize_static")
    for k in ["cellType", "svx", "svy", "svz"]:
        assert np.array_equal(st[k], g["rast_" + k]), k
    assert int(((g["rast_cellType"] == 2) & (g["static_types"] != 2)).sum()) > 50
    st.call("oc_mark_fluid", 0, st.get_i("nP"))
    assert np.array_equal(st["cellType"], g["mark_cellType"])


def test_ball_push_out_bit_exact():
    g = golden.load("balls_32")
    st = state_with_balls(g)
    st.call("oc_clear_faces")
    st.call("oc_rasterize_static")
    st.call("oc_mark_fluid", 0, st.get_i("nP"))
    st.call("oc_push_balls")
    moved = 0
    for k in ["px", "py", "pz", "pvx", "pvy", "pvz"]:
        assert np.array_equal(st[k][:len(g["pushb_" + k])], g["pushb_" + k]), k
        moved += int((g["pushb_" + k] != g["in_" + k]).sum())
    assert moved > 20


def test_substep_with_balls_bit_exact():
    """One whole substep: the solid velocities enter the boundary conditions,
    the pressure and the particles; the integrator's histogram, waterlines and
    every ball's nine numbers match."""
    g = golden.load("balls_32")
    st = state_with_balls(g)
    st.call("oc_substep", g.meta["stage_dt"])
    n = len(g["adv_px"])
    for k in PART_KEYS:
        assert np.array_equal(st[k][:n], g["adv_" + k]), k
    hist, surf = st.ball_hist()
    assert np.array_equal(hist.ravel(), g["ball_hist"])
    assert np.array_equal(surf, g["ball_surf"])
    assert np.array_equal(st.balls().ravel(), g["out_balls"])


def test_ball_trajectory_replay_bit_exact():
    g = golden.load("balls_32")
    st = state_with_balls(g, tag="fill_")
    st.arrays["q"][:] = g["fill_q"]
    st.set_params(gRamp=g.meta["fill"]["gRamp"])
    st.set_f("umax", 0.0)
    rec = g.meta["trajectory"]
    tb = g["traj_balls"].reshape(len(rec), -1, 9)
    for i in range(len(rec)):
        dt = st.call("oc_dt_law")
        assert dt == rec[i]["dt"], i
        st.call("oc_substep", dt)
        assert st.get_f("umax") == rec[i]["umax"], i
        assert np.array_equal(st.balls(), tb[i]), i
    n = len(g["traj_end_px"])
    for k in PART_KEYS:
        assert np.array_equal(st[k][:n], g["traj_end_" + k]), k


def test_balls_move_and_touch():
    """The scene exercises the branches it was built for: the dropped ball
    hits the water, the pair separates, the land ball is pushed out of the
    rock and rolls."""
    g = golden.load("balls_32")
    b0 = g["fill_balls"].reshape(-1, 9)
    b1 = g["traj_balls"].reshape(20, -1, 9)[-1]
    assert b1[1, 8] > 0.3            # dropped ball submerged
    d0 = np.linalg.norm(b0[2, :3] - b0[3, :3])
    d1 = np.linalg.norm(b1[2, :3] - b1[3, :3])
    assert d0 < b0[2, 6] + b0[3, 6] <= d1 + 1e-3    # overlap resolved
    assert np.linalg.norm(b1[4, :3] - b0[4, :3]) > 0.05  # land ball moved
