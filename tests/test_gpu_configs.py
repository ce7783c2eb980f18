"""The benchmark configurations on the GPU at their own sizes, against the
CPU oracle, plus the particle-to-cell index at float boundaries.

- cfg2 (64^3, 220k particles, curl-noise start, vorticity, currents) and
  cfg3 (128^3, 2M particles, full physics): the scene bench.py times, three
  device-dt substeps in, then one more substep on the GPU and in the oracle
  from the same state with the same dt and the same number of red-black
  sweeps (the GPU's last solve count). The marking pass is checked first on
  its own: cell types and flags bit-exact, and every marked particle's cell
  index equal to clamp(floor(double(x)/dx)) of its float coordinates
  (js/solver.js:976-978). After the substep: cell types bit-exact, particles
  within the single-substep bar (2e-5 of the field scale, at most 0.1 % on
  another discrete branch), faces within 1e-4. These sizes run the 26-bit
  list codes, the 2M-particle counting sorts and the large P2G tiles.
- The particle-to-cell index is also pinned on float coordinates placed on
  both sides of every cell boundary along each axis (the smallest float
  whose double quotient floors to i, and the float just below it), where a
  float division would round the wrong way.
- cfg3 with the temporally blocked SOR (two sweeps per launch) gives the
  bits of the fluid-list sweeps.
- cfg4: the shipped network weights on the held-out benchmark seed: the
  learned start reaches residual 1e-4 in fewer sweeps than the previous
  pressure on the same systems; the time ratio is printed.
- cfg5: a batch that includes a 96^3 planet equals the planets run one by
  one (bit for bit, learned warm start on), and its 32^3 planet's batched
  substep matches the oracle's substep from the same state. The batch solves
  from the network's guess to residual 1e-4; the oracle runs the substep
  twice, to 1e-4 (A) and to 1e-7 (B). The batch is held to three times the
  spread |A - B| between two solutions the stopping rule accepts (plus the
  2e-5 float bar), measured from B.
"""
import ctypes
import os

import numpy as np
import pytest

import bench
from gpu_scenes import oracle_from_solver, watersim_hip

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "watersim-v1_amd", "weights", "warmstart_unet.f32")
PK = ["px", "py", "pz", "pvx", "pvy", "pvz", "pflag", "pcool", "pAir"]


def floor_index(x, dx, n):
    return np.clip(np.floor(x.astype(np.float64) / dx), 0, n - 1).astype(np.int64)


def compare_particles(got, ref, n, tol, what, max_flip=0.001, pos_tol=None):
    bad = np.zeros(n, bool)
    for k in ("pflag", "pcool"):
        bad |= got[k] != ref[k][:n]
    errs = {}
    for k in ("px", "py", "pz", "pvx", "pvy", "pvz", "pAir"):
        t = pos_tol if (pos_tol is not None and k in ("px", "py", "pz")) else tol
        d = np.abs(got[k].astype(np.float64) - ref[k][:n])
        bad |= d > t * max(1.0, float(np.abs(ref[k][:n]).max()))
    for k in ("px", "py", "pz", "pvx", "pvy", "pvz"):
        d = np.abs(got[k].astype(np.float64) - ref[k][:n])
        errs[k] = float(d[~bad].max()) if (~bad).any() else 0.0
    print("  %s: %d / %d on another branch or off, max |err| on the rest %s" % (what, int(bad.sum()), n, errs))
    assert bad.sum() <= max(1, int(max_flip * n)), (what, int(bad.sum()))


@pytest.mark.parametrize("name", ["cfg2", "cfg3"])
def test_full_size_substep_vs_oracle(name):
    cfg = dict(bench.CONFIGS[name])
    s, world = bench.make_solver(cfg)
    n = cfg["n"]
    s.run_substeps(3)
    sweeps = int(s.get("pressureIters_done"))
    nP = int(s.get("nP"))
    o = oracle_from_solver(s, world)
    print("\n[%s] %d^3, %d particles, %d sweeps per solve" % (name, n, nP, sweeps))
    # the marking pass alone
    x, y, z = s.download("px"), s.download("py"), s.download("pz")
    s.stage("rasterize", 0.0)
    o.call("oc_rasterize_static")
    o.call("oc_mark_fluid", 0, nP)
    assert np.array_equal(s.download("cellType"), o["cellType"])
    assert np.array_equal(s.download("pflag"), o["pflag"][:nP])
    pc = s.download("pcell")
    marked = o["pflag"][:nP] == 0
    want = (floor_index(z, world["dx"], n) * n + floor_index(y, world["dx"], n)) * n + floor_index(x, world["dx"], n)
    assert np.array_equal(pc[marked], want[marked])
    assert np.all(pc[~marked] == -1)
    print("  mark: cell types, flags and %d particle cell indices exact" % int(marked.sum()))
    # one substep on both, same dt, same sweeps
    s.set(pressureTol=0, pressureIters=sweeps)
    o.set_params(pressureIters=sweeps, sorRedBlack=1)
    dt = o.call("oc_dt_law")
    s.substep(dt)
    o.call("oc_substep", ctypes.c_double(dt))
    assert np.array_equal(s.download("cellType"), o["cellType"])
    got = {k: s.download(k) for k in PK}
    compare_particles(got, o.arrays, nP, 2e-5, "substep particles")
    for k in "uvw":
        a, b = s.download(k).astype(np.float64), o[k].astype(np.float64)
        e = np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
        print("  face %s max rel err %.3e" % (k, e))
        assert e <= 1e-4, (k, e)
    assert abs(s.get("umax") - o.get_f("umax")) <= 1e-4 * o.get_f("umax")
    s.close()


@pytest.mark.skip(reason="the row-pipelined extrapolation is switched off: its first GPU run (a 64^3 "
                         "substep) stalled and no cause was found by reading the code")
def test_row_extrapolation_at_128():
    """The row-pipelined extrapolation (exRows 1: flags as a 3-D prefix OR
    of the seeds, values in one launch of slabs x chunks that hand rows over
    through one XCD's L2) against the tile-diagonal kernel (exRows 0) and
    the oracle's serial flood (js/solver.js:1562-1636) on the cfg3 scene at
    its own size, three substeps in: every final flag exact, values within
    2e-5 of the field scale (the reference sums in double, these kernels
    scan affine maps in float)."""
    cfg = bench.CONFIGS["cfg3"]
    s, world = bench.make_solver(cfg)
    s.run_substeps(3)
    keys = ("u", "v", "w", "uW", "vW", "wW")
    keep = {k: s.download(k) for k in keys}
    res = {}
    for rows in (1, 0, 1):
        for k in keys:
            s.upload(k, keep[k])
        s.set(exRows=rows)
        s.stage("extrapolate")
        out = {k: s.download(k) for k in ("u", "v", "w", "validU", "validV", "validW")}
        if rows in res:  # the second row-pipelined run repeats the first bit for bit
            for k in out:
                assert np.array_equal(out[k], res[rows][k]), ("repeat", k)
        res[rows] = out
    o = oracle_from_solver(s, world)
    o.arrays["cellType"][:] = s.download("cellType")
    for k in keys:
        o.arrays[k][:] = keep[k]
    o.call("oc_extrapolate")
    filled = 0
    for k, vk in (("u", "validU"), ("v", "validV"), ("w", "validW")):
        assert np.array_equal(res[1][vk], o[vk]), vk
        assert np.array_equal(res[0][vk], o[vk]), vk
        scale = float(np.abs(o[k]).max())
        for rows in (1, 0):
            e = float(np.abs(res[rows][k].astype(np.float64) - o[k]).max()) / scale
            print("  %s exRows %d: max err %.3e of the field scale" % (k, rows, e))
            assert e <= 2e-5, (k, rows, e)
        filled += int(np.sum(res[1][vk] != (keep[k + "W"] > 1e-8)))
    print("  flags exact; %d faces valid without a P2G weight" % filled)
    s.close()


def test_cell_index_at_float_boundaries():
    """Particles on both sides of every cell boundary of each axis, the other
    two coordinates at the domain centre. The planet's demotion rules are
    switched off (ocean radius and terrain far beyond the domain), so every
    particle marks a cell and reports its index."""
    n, W = 32, 5.1
    dx = W / n
    xs = []
    for i in range(n + 1):
        f = np.float32(i * dx)
        # the smallest float whose double quotient floors to at least i
        while np.floor(np.float64(f) / dx) >= i and f > 0:
            f = np.nextafter(f, np.float32(-1))
        while np.floor(np.float64(f) / dx) < i:
            f = np.nextafter(f, np.float32(1e9))
        xs += [f, np.nextafter(f, np.float32(-1)), np.nextafter(f, np.float32(1e9))]
    xs = np.array([v for v in xs if 0 <= v <= W], np.float32)
    c = np.float32(0.5 * W)
    pts = []
    for a in range(3):
        p = np.full((3, len(xs)), c, np.float32)
        p[a] = xs
        pts.append(p)
    p = np.concatenate(pts, axis=1)
    N = p.shape[1]
    s = watersim_hip.Solver(n, n, n, dx, N)
    s.set(coreR=0.25 * W, oceanR=100.0, spacing=dx / 2, particleVolume=(dx / 2) ** 3, dv=dx / 2, tn=2 * n)
    s.upload("staticTypes", np.zeros(n ** 3, np.uint8))
    for k, m in (("gu", (n + 1) * n * n), ("gv", (n + 1) * n * n), ("gw", (n + 1) * n * n)):
        s.upload(k, np.zeros(m, np.float32))
    s.upload("terrainR", np.full((2 * n) ** 3, 100.0, np.float32))
    s.set(nP=N)
    for k, v in zip(("px", "py", "pz"), p):
        s.upload(k, v)
    s.upload("pflag", np.zeros(N, np.uint8))
    s.stage("rasterize", 0.0)
    assert np.all(s.download("pflag") == 0)
    pc = s.download("pcell")
    want = (floor_index(p[2], dx, n) * n + floor_index(p[1], dx, n)) * n + floor_index(p[0], dx, n)
    # a float division would differ somewhere here; the cell index may not
    naive = np.clip(np.floor(p.astype(np.float32) / np.float32(dx)), 0, n - 1).astype(np.int64)
    naive_idx = (naive[2] * n + naive[1]) * n + naive[0]
    print("  %d boundary coordinates; a float32 division would misplace %d of them"
          % (N, int((naive_idx != want).sum())))
    assert np.array_equal(pc, want)
    s.close()


def test_learned_warm_start_gain():
    """cfg4: the shipped weights on the benchmark scene (its seed is held
    out of training), every solve run from the previous pressure and from
    the network's guess to the same residual."""
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained warm-start weights")
    cfg = dict(bench.CONFIGS["cfg4"])
    s, world = bench.make_solver(cfg)
    s.load_net_weights(WEIGHTS)
    bench.advance(s, world, cfg, 6, 0)
    s.sync()
    s.set(timing=1, pressureCompare=1)
    bench.advance(s, world, cfg, 6, 6)
    s.sync()
    g = {k: s.get(k) for k in ("cmpMsPlain", "cmpMsLearned", "cmpSweepsPlain", "cmpSweepsLearned", "cmpSolves")}
    s.set(timing=0, pressureCompare=0)
    ns = g["cmpSolves"]
    print("\n  %d solves: sweeps plain %.1f learned %.1f; ms plain %.3f learned %.3f (speed-up %.2fx)"
          % (ns, g["cmpSweepsPlain"] / ns, g["cmpSweepsLearned"] / ns, g["cmpMsPlain"] / ns,
             g["cmpMsLearned"] / ns, g["cmpMsPlain"] / g["cmpMsLearned"]))
    assert ns >= 6
    assert g["cmpSweepsLearned"] < g["cmpSweepsPlain"]
    s.close()


def test_temporal_blocking_equals_list_sweeps_at_128():
    """The two-sweeps-per-launch temporally blocked SOR (sorMode 3: 32x16x16
    tiles, per-tile phase lists) against the fluid-list kernel on the cfg3
    scene at its own size: every substep's solve to 1e-4 (an odd sweep
    count ends with a one-sweep launch), the same bits in q and every
    face and particle."""
    cfg = bench.CONFIGS["cfg3"]
    out = []
    for mode in (2, 3):
        s, world = bench.make_solver(cfg)
        s.set(sorMode=mode)
        s.run_substeps(4)
        out.append({k: s.download(k).copy() for k in ("q", "u", "v", "w", "px", "pvy")})
        out[-1]["sweeps"] = s.get("sweepsTotal")
        s.close()
    a, b = out
    assert a["sweeps"] == b["sweeps"] and a["sweeps"] > 0
    for k in ("q", "u", "v", "w", "px", "pvy"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("groups,units", [(0, 0), (31, 0), (32, 3)])
def test_xcd_sweeps_equal_list_sweeps_at_128(groups, units):
    """The XCD-resident SOR (sorMode 6: every batch of sweeps between two
    residual checks in one launch on the workgroups of one XCD, q of each
    workgroup's cell range in LDS, neighbour ranges exchanged through L2)
    against the fluid-list kernel on the cfg3 scene at its own size, for
    several workgroup counts (0: one per compute unit of the XCD): every
    substep's solve to 1e-4, the same sweep counts and the same bits in q,
    every face and every particle. No launch may decline."""
    cfg = bench.CONFIGS["cfg3"]
    out = []
    for mode in (2, 6):
        s, world = bench.make_solver(cfg)
        s.set(sorMode=mode, xcdGroups=groups, xcdUnits=units)
        s.run_substeps(4)
        out.append({k: s.download(k).copy() for k in ("q", "u", "v", "w", "px", "pvy")})
        out[-1]["sweeps"] = s.get("sweepsTotal")
        out[-1]["fallbacks"] = s.get("xcdFallbacks")
        out[-1]["mode"] = s.get("sorMode")
        s.close()
    a, b = out
    assert b["fallbacks"] == 0 and b["mode"] == 6
    assert a["sweeps"] == b["sweeps"] and a["sweeps"] > 0
    for k in ("q", "u", "v", "w", "px", "pvy"):
        assert np.array_equal(a[k], b[k]), k


def test_xcd_sweeps_decline_to_list_kernel():
    """A launch whose cell ranges do not fit its workgroups' LDS (forced from
    the host: one cell per thread and colour at 128^3) declines before it
    writes anything; the host sees it at the residual check, runs the same
    sweeps with the list kernel and, the size being pinned by the setting,
    keeps that kernel. The results are the list kernel's bits."""
    cfg = bench.CONFIGS["cfg3"]
    out = []
    for mode in (2, 6):
        s, world = bench.make_solver(cfg)
        s.set(sorMode=mode, xcdUnits=1)
        s.run_substeps(2)
        out.append({k: s.download(k).copy() for k in ("q", "u", "px")})
        out[-1]["sweeps"] = s.get("sweepsTotal")
        out[-1]["fallbacks"] = s.get("xcdFallbacks")
        out[-1]["mode"] = s.get("sorMode")
        s.close()
    a, b = out
    assert b["fallbacks"] == 1 and b["mode"] == 2
    assert a["sweeps"] == b["sweeps"]
    for k in ("q", "u", "px"):
        assert np.array_equal(a[k], b[k]), k


def _planet(n, seed):
    w = watersim_hip.build_world(n, bench.W_DOMAIN / n, 0.25 * bench.W_DOMAIN, 0.15 * bench.W_DOMAIN,
                                 20260500 + seed, int(0.8 * n ** 3))
    s = watersim_hip.solver_for_world(w)
    s.set(pressureTol=1e-4, maxPressureIters=4000, residualEvery=8, vorticity=0.05, cfl=3.0, maxSpeed=3.2,
          pic=0.05, omega=1.7, gravity=9.8)
    s.load_net_weights(WEIGHTS)
    s.set(warmStart=1)
    return s, w


def test_batch_with_96_planet():
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained warm-start weights")
    spec = [(32, 1), (96, 2), (64, 3)]
    ref = [_planet(n, sd) for n, sd in spec]
    bat = [_planet(n, sd) for n, sd in spec]
    for s, _ in ref:
        s.set(sorMode=4, sorThreads=512)
    b = watersim_hip.Batch([s for s, _ in bat], graph=True, threads=512)
    b.run(2)
    for s, _ in ref:
        s.run_substeps(2)
    for i, ((r, _), (s, _)) in enumerate(zip(ref, bat)):
        for k in ("q", "u", "px", "pvx", "pflag", "cellType"):
            assert np.array_equal(r.download(k), s.download(k)), (i, k)
    print("\n  batch of 32^3, 96^3, 64^3: every planet equals its single run after 2 substeps")
    # the 32^3 planet's next batched substep against the oracle's, run from
    # the same state to residual 1e-4 (A) and to 1e-7 (B): two solutions the
    # stopping rule accepts differ by |A - B|, so the batch (solved from the
    # network's guess to 1e-4) is held to 3 |A - B| from B
    s32, w32 = bat[0]
    nP = int(s32.get("nP"))
    oa, ob = oracle_from_solver(s32, w32), oracle_from_solver(s32, w32)
    oa.set_params(pressureTol=1e-4, maxPressureIters=4000, residualEvery=8, sorRedBlack=1)
    ob.set_params(pressureTol=1e-7, maxPressureIters=20000, residualEvery=8, sorRedBlack=1)
    dt = oa.call("oc_dt_law")
    b.run(1)
    assert s32.get("dtLast") == np.float32(dt)
    for o in (oa, ob):
        o.call("oc_substep", ctypes.c_double(dt))
    assert np.array_equal(s32.download("cellType"), ob["cellType"])
    same = (s32.download("pflag") == ob["pflag"][:nP]) & (oa["pflag"][:nP] == ob["pflag"][:nP])
    same &= (s32.download("pcool") == ob["pcool"][:nP]) & (oa["pcool"][:nP] == ob["pcool"][:nP])
    print("  batched 32^3 substep: %d / %d particles on another branch" % (int((~same).sum()), nP))
    assert (~same).sum() <= max(1, int(0.001 * nP))
    for k in ("px", "py", "pz", "pvx", "pvy", "pvz"):
        ref_b = ob[k][:nP].astype(np.float64)[same]
        spread = float(np.abs(oa[k][:nP].astype(np.float64)[same] - ref_b).max())
        err = float(np.abs(s32.download(k).astype(np.float64)[same] - ref_b).max())
        print("    %-4s |batch - B| %.3e, |A - B| %.3e" % (k, err, spread))
        assert err <= 3 * spread + 2e-5 * max(1.0, float(np.abs(ref_b).max())), (k, err, spread)
    b.close()
    for s, _ in ref + bat:
        s.close()
