"""The drivers above the substep on the GPU, against the reference's own
trajectory and the oracle's frame loop:

- ws_run_substeps evaluates the CFL law of _stepSync (js/solver.js:2689-2690)
  on the device. From the cfg1 fill state it runs the 100 substeps the
  reference recorded (tests/golden/gen_golden.js `trajectory`): every dt
  equals the reference's (as the float the kernels use), the gravity ramp
  advances by exactly those dts, umax and the kinetic energy follow the
  reference. The reference sweeps lexicographic SOR, the GPU red-black SOR
  (same 40 sweeps), so the bars are those the oracle itself shows between
  the two orderings over these 100 substeps (KE 4.8e-4, umax 5.9e-7
  relative), with room for FP32 summation order: KE 2e-3, umax 1e-4
  relative. Against the oracle's red-black run the bars are KE 1e-3 and
  umax 1e-4.
- ws_step_frame is _stepSync: the 25 Hz heat / evaporation / currents tick,
  the substep loop, then vapor collisions, phase changes, snow contacts and
  the surface rescue in that order (:2670-2700). It is run for six 1/60 s
  frames from the thermal scene's tick input state, beside the same frames
  driven over the oracle's stages (tests/scenes.py oracle_step_frame): the
  substep count of every frame, simTime and heatAcc (host doubles) and the
  tick count are exact, every frame's last dt equals the oracle's as a
  float, and the evaporation count stays within 1 % (one candidate whose
  draw lands on a float tie of its threshold moves the stream). Particles
  are compared after the six frames (12 substeps, two ticks, six contact
  passes): a particle that took another discrete branch on a float tie
  (flag, cooldown) or ended beyond 1e-4 counts as off, at most 1 % of them
  (a flip moves its neighbours through the grid and the contacts); the rest
  agree within 1e-3.
"""
import numpy as np
import pytest

import golden
from gpu_scenes import gpu_from_golden
from scenes import oracle_from_golden, oracle_step_frame, PART_KEYS

pytestmark = pytest.mark.gpu


def test_device_dt_law_trajectory_vs_reference():
    g = golden.load("cfg1_32")
    traj = g.meta["trajectory"]
    s = gpu_from_golden(g, tag="fill_")
    g0 = g.meta["fill"]["gRamp"]
    s.set(gRamp=g0)
    o = oracle_from_golden(g, tag="fill_")
    o.set_params(gRamp=g0, sorRedBlack=1)
    ramp = g0
    worst = {"ke_ref": 0.0, "um_ref": 0.0, "ke_orc": 0.0, "um_orc": 0.0}
    for i, rec in enumerate(traj):
        s.run_substeps(1)
        dt_o = o.call("oc_dt_law")
        o.call("oc_substep", dt_o)
        assert dt_o == rec["dt"], (i, dt_o, rec["dt"])
        assert s.get("dtLast") == np.float32(rec["dt"]), (i, s.get("dtLast"), rec["dt"])
        ramp += rec["dt"]
        assert s.get("gRamp") == ramp, (i, s.get("gRamp"), ramp)
        ke, um = s.kinetic_energy(), s.get("umax")
        ke_o, um_o = o.call("oc_kinetic_energy"), o.get_f("umax")
        if rec["ke"] > 0:
            worst["ke_ref"] = max(worst["ke_ref"], abs(ke - rec["ke"]) / rec["ke"])
            worst["um_ref"] = max(worst["um_ref"], abs(um - rec["umax"]) / rec["umax"])
            worst["ke_orc"] = max(worst["ke_orc"], abs(ke - ke_o) / ke_o)
            worst["um_orc"] = max(worst["um_orc"], abs(um - um_o) / um_o)
        if i % 20 == 19:
            print("  substep %3d ke %.6e (reference %.6e) umax %.6e (reference %.6e)"
                  % (i + 1, ke, rec["ke"], um, rec["umax"]))
    print("  worst relative errors over 100 substeps: %s" % worst)
    assert worst["ke_ref"] <= 2e-3 and worst["um_ref"] <= 1e-4, worst
    assert worst["ke_orc"] <= 1e-3 and worst["um_orc"] <= 1e-4, worst
    s.close()


def _frame_pair():
    """The thermal scene at its tick input state on both sides, with the
    frame clock set so that the first frame runs a tick."""
    g = golden.load("thermal_32")
    s = gpu_from_golden(g, tag="t0_")
    o = oracle_from_golden(g, tag="t0_")
    for k in ("cellType", "dens", "u", "v", "w"):
        s.upload(k, g["t0_" + k])
        o.arrays[k][:] = g["t0_" + k]
    t = g.meta["tick"]
    rb = g.meta["rng_before"]
    s.set_rng(rb["cfS"], rb["bfS"], t["evS_before"])
    o.set_params(cfS=rb["cfS"], bfS=rb["bfS"], evS=t["evS_before"])
    sp = g.meta["frame_end"]["sunPos"]
    s.set(spinOn=1, hasVapor=1, sunX=sp[0], sunY=sp[1], sunZ=sp[2], simTime=t["simTime"], heatAcc=0.03)
    # a fresh context's umax is 0, so the first dt comes from the 0.08 floor
    # of the law on both sides
    o.set_params(spinOn=1, sunX=sp[0], sunY=sp[1], sunZ=sp[2], umax=0.0)
    s.set(gRamp=g.meta["stage_params"]["gRamp"])
    o.set_params(gRamp=g.meta["stage_params"]["gRamp"])
    return g, s, o, {"simTime": t["simTime"], "heatAcc": 0.03}


def test_step_frame_vs_oracle_frames():
    g, s, o, clock = _frame_pair()
    t = g.meta["tick"]
    n = len(g["t0_px"])
    ticks0 = s.get("ticksTotal")
    for f in range(6):
        nsub = s.step(1.0 / 60)
        dts = oracle_step_frame(o, 1.0 / 60, clock, t["sunActivity"], t["currents"])
        print("  frame %d: substeps %d / %d, dt %s / last %.6g, evap %d / %d, vapor %d / %d"
              % (f, nsub, len(dts), ["%.6g" % d for d in dts], s.get("dtLast"), s.get("evapCount"),
                 o.get_i("evapCount"), s.get("vaporCount"), o.get_i("vaporCount")))
        assert nsub == len(dts)
        assert s.get("dtLast") == np.float32(dts[-1])
        assert s.get("simTime") == clock["simTime"]
        assert s.get("heatAcc") == clock["heatAcc"]
        assert s.get("ticksTotal") - ticks0 == clock.get("ticks", 0)
        assert abs(s.get("evapCount") - o.get_i("evapCount")) <= max(1, 0.01 * o.get_i("evapCount"))
    got = {k: s.download(k) for k in PART_KEYS}
    ref = {k: o[k][:n] for k in PART_KEYS}
    bad = np.zeros(n, bool)
    for k in ("pflag", "pcool"):
        bad |= got[k] != ref[k]
    errs = {}
    for k in ("px", "py", "pz", "pvx", "pvy", "pvz", "pT"):
        d = np.abs(got[k].astype(np.float64) - ref[k])
        scale = max(1.0, float(np.abs(ref[k]).max()))
        bad |= d > 1e-4 * scale
    for k in ("px", "py", "pz", "pvx", "pvy", "pvz", "pT"):
        errs[k] = float(np.abs(got[k].astype(np.float64) - ref[k])[~bad].max())
    print("  after 6 frames: %d / %d particles off (flag flips or beyond 1e-4), max |err| on the rest %s"
          % (int(bad.sum()), n, errs))
    assert bad.sum() <= max(1, int(0.01 * n))
    assert max(errs.values()) <= 1e-3, errs
    s.close()
