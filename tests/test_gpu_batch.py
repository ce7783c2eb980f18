"""The one-workgroup pressure solve (sorMode 4) and batched planets
(ws_batch_*); neither has a counterpart in the reference, which runs one
solver with a host-side SOR loop.

- sorMode 4 runs the sweeps of the fluid-list kernel (sorMode 2) and the
  host loop's stopping rule on the device (sorMode 5 the same on workgroups
  synchronised with their neighbours only). The sweeps give the same bits and
  the stopping decisions agree, so whole substeps agree bit for bit: with a
  fixed sweep count, to a residual with checks every 8 sweeps, and with the
  adaptive check spacing (same sweep counts; the residual itself is summed in
  another order, so it is held to 1e-12 relative).
- A batch of planets of mixed sizes (one with the learned warm start) gives,
  for every planet, exactly what ws_run_substeps gives with sorMode 4: run
  eagerly, and replayed from its HIP graph, with each solve in one
  workgroup or split over many (the residual is summed in an order fixed by
  the list, not by the split). The graph path is called several
  times, with downloads and synchronizations between the calls, and every
  call's results are checked.
- The batched-planets workload itself, 64 planets of 32^3..96^3, runs from
  its graph with every planet at the tolerance, and one planet of each size
  equals its own single-planet run.
- The warm start's forward pass, batched over the planets, gives each
  planet the bits of its own pass in all three forms of the convolution.
"""
import os

import numpy as np
import pytest

from gpu_scenes import watersim_hip

pytestmark = pytest.mark.gpu

W = 5.1
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "watersim-v1_amd", "weights", "warmstart_unet.f32")
FIELDS = ("q", "u", "v", "w", "px", "py", "pz", "pvx", "pvy", "pvz", "cellType")


def planet(n, seed, warm=False):
    w = watersim_hip.build_world(n, W / n, 0.25 * W, 0.15 * W, 20260400 + seed, int(0.8 * n ** 3))
    s = watersim_hip.solver_for_world(w)
    s.set(pressureTol=1e-4, maxPressureIters=4000, residualEvery=8, vorticity=0.05, cfl=3.0, maxSpeed=3.2)
    if warm:
        s.load_net_weights(WEIGHTS)
        s.set(warmStart=1)
    return s


def assert_same(a, b, what):
    for k in FIELDS:
        x, y = a.download(k), b.download(k)
        bad = int(np.sum(x != y))
        assert bad == 0, "%s: %s differs in %d entries" % (what, k, bad)


@pytest.mark.parametrize("adaptive", [0, 1])
def test_device_solve_equals_list_sweeps(adaptive):
    a, b = planet(32, 1), planet(32, 1)
    for s, mode in ((a, 2), (b, 4)):
        s.set(adaptiveChecks=adaptive, sorMode=mode)
    for step in range(4):
        a.run_substeps(1)
        b.run_substeps(1)
        na, nb = a.get("pressureIters_done"), b.get("pressureIters_done")
        ra, rb = a.get("residual"), b.get("residual")
        print("  substep %d: sweeps %d / %d, residual %.6g / %.6g" % (step, na, nb, ra, rb))
        assert na == nb
        assert abs(ra - rb) <= 1e-12 * ra
        assert_same(a, b, "substep %d" % step)
    assert a.get("sweepsTotal") == b.get("sweepsTotal")
    a.close()
    b.close()


@pytest.mark.parametrize("groups", [4, 16])
def test_device_solve_split_over_workgroups(groups):
    """One context's solve over several workgroups meeting at barriers gives
    the bits of the list sweeps (and of the one-workgroup solve)."""
    a, b = planet(48, 4), planet(48, 4)
    a.set(sorMode=2)
    b.set(sorMode=4, sorGroups=groups)
    for step in range(3):
        a.run_substeps(1)
        b.run_substeps(1)
        assert a.get("pressureIters_done") == b.get("pressureIters_done")
        assert_same(a, b, "substep %d" % step)
    a.close()
    b.close()


@pytest.mark.parametrize("groups,threads", [(256, 256), (24, 512)])
def test_flow_solve_equals_list_sweeps(groups, threads):
    """The neighbour-synchronised solve (sorMode 5: workgroups own ranges of
    the lists and wait only for the workgroups within one plane of cells)
    gives the bits, sweep counts and residuals of the list sweeps; 24
    workgroups of 512 threads make each range span several planes."""
    a, b = planet(48, 5), planet(48, 5)
    a.set(sorMode=2)
    b.set(sorMode=5, flowGroups=groups, sorThreads=threads)
    for step in range(3):
        a.run_substeps(1)
        b.run_substeps(1)
        assert a.get("pressureIters_done") == b.get("pressureIters_done")
        assert abs(a.get("residual") - b.get("residual")) <= 1e-12 * a.get("residual")
        assert_same(a, b, "substep %d" % step)
    for s in (a, b):
        s.set(pressureTol=0, pressureIters=37, measureResidual=1)
        s.run_substeps(1)
    assert b.get("pressureIters_done") == 37
    assert_same(a, b, "fixed sweeps")
    a.close()
    b.close()


def test_device_solve_fixed_sweeps():
    a, b = planet(32, 2), planet(32, 2)
    for s, mode in ((a, 2), (b, 4)):
        s.set(pressureTol=0, pressureIters=37, measureResidual=1, sorMode=mode)
        s.run_substeps(2)
    assert b.get("pressureIters_done") == 37
    assert abs(a.get("residual") - b.get("residual")) <= 1e-12 * a.get("residual")
    assert_same(a, b, "fixed sweeps")
    a.close()
    b.close()


@pytest.mark.parametrize("graph,cells_per_group,xcd_local,max_groups,local_planets", [
    (False, 110000, 1, 256, 1), (True, 110000, 1, 256, 1), (True, 4000, 1, 256, 3), (True, 4000, 0, 256, 0),
    (True, 4000, 1, 96, 2), (True, 4000, 1, 40, 0)])
def test_batch_equals_single_planets(graph, cells_per_group, xcd_local, max_groups, local_planets):
    """cells_per_group 110000 gives the 48^3 planet two workgroups; 4000
    splits the solves over 9, 28 and 9 workgroups. With xcd_local each
    planet's workgroups are laid out on one XCD where they fit (32 slots per
    XCD at max_groups 256: all three planets; 12 at 96: the 9-workgroup
    planets only; 5 at 40, where the 48^3 planet is cut to 22: none) and take
    the L2-local barriers; the others take the agent-scope ones. Either way
    the bits are those of ws_run_substeps."""
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained warm-start weights")
    spec = [(32, 11, False), (48, 12, True), (32, 13, False)]
    ref = [planet(n, sd, wm) for n, sd, wm in spec]
    bat = [planet(n, sd, wm) for n, sd, wm in spec]
    for s in ref:
        s.set(sorMode=4, sorThreads=512)
    b = watersim_hip.Batch(bat, graph=graph, threads=512)
    b.set(cellsPerGroup=cells_per_group, xcdLocal=xcd_local, maxGroups=max_groups)
    for call, k in enumerate((1, 2, 3)):
        b.run(k)
        for s in ref:
            s.run_substeps(k)
        for i, (r, s) in enumerate(zip(ref, bat)):
            assert_same(r, s, "call %d planet %d" % (call, i))
            assert r.get("sweepsTotal") == s.get("sweepsTotal")
            assert r.get("residual") == s.get("residual")
        print("  call %d (%d substeps): %d planets equal, sweeps %s, %d blocks, %d planets XCD-local" % (
            call, k, len(bat), [int(s.get("sweepsTotal")) for s in bat], int(b.get("groups")),
            int(b.get("xcdLocalPlanets"))))
        assert int(b.get("xcdLocalPlanets")) == local_planets
    b.close()
    for s in ref + bat:
        s.close()


@pytest.mark.parametrize("net_batch,stage_batch", [(1, 1), (1, 0), (0, 0)])
def test_batched_network_pass_equals_planets_own(net_batch, stage_batch):
    """The warm start's forward pass and correction run as one launch per
    network step over every planet with the warm start on (netBatch 1) or
    as each planet's own launches (0). The planets cover the three forms of
    the convolution (64^3 at full resolution: LDS-tiled at its finest level,
    four waves per 16 voxels at the coarser ones; a 32^3 grid coarsened to
    16^3 with netTiled 2: LDS-tiled down to 4^3; 48^3 with netTiled 0:
    untiled) and both corrections (coarsened and full), beside a
    planet without the warm start. The planets' own stages run merged into
    one launch per stage (stageBatch 1, which batches the network's pass
    too; the planets differ in grid size, so in viscosity iterations, and in
    the network's input stage) or on their own streams (0). Either way every planet's state equals its own
    ws_run_substeps run bit for bit, eagerly and from the graph."""
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained warm-start weights")
    spec = [(32, 21, True, dict(netCoarse=1, netTiled=2)), (64, 22, True, dict(netCoarse=0)),
            (48, 23, True, dict(netCoarse=0, netTiled=0)), (32, 24, False, {})]

    def make():
        out = []
        for n, sd, wm, extra in spec:
            s = planet(n, sd, wm)
            if extra:
                s.set(**extra)
            out.append(s)
        return out
    ref, bat = make(), make()
    for s in ref:
        s.set(sorMode=4, sorThreads=512)
    for graph in (False, True):
        b = watersim_hip.Batch(bat, graph=graph, threads=512)
        b.set(netBatch=net_batch, stageBatch=stage_batch)
        assert int(b.get("netBatch")) == net_batch
        b.run(2)
        launches = int(b.get("stageLaunches"))
        assert (launches > 0) == bool(stage_batch)
        for s in ref:
            s.run_substeps(2)
        for i, (r, s) in enumerate(zip(ref, bat)):
            assert_same(r, s, "graph %d planet %d" % (graph, i))
            assert r.get("sweepsTotal") == s.get("sweepsTotal")
        b.close()
    print("  sweeps %s, %d stage launches per batched substep" % ([int(s.get("sweepsTotal")) for s in bat], launches))
    for s in ref + bat:
        s.close()


def test_batch_of_64_planets_as_benchmarked():
    """The batched-planets workload at its own size (the benchmark's draw of
    64 planets with grid sides 32, 48, 64, 96, about 8 particles per liquid
    cell, own terrain seed and gravity, learned warm start, SOR to 1e-4),
    two substeps replayed from the batch's HIP graph. Every planet ends at
    the tolerance with finite state inside the domain, and the first planet
    of each size equals its own run through ws_run_substeps (sorMode 4) bit
    for bit."""
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained warm-start weights")
    import bench
    cfg = bench.CONFIGS["cfg5"]
    planets, sizes = bench.make_planets(cfg, WEIGHTS)
    assert len(planets) == 64 and sorted(set(sizes)) == [32, 48, 64, 96]
    firsts = [sizes.index(n) for n in (32, 48, 64, 96)]
    ref, _ = bench.make_planets(cfg, WEIGHTS, only=firsts)
    for s in ref:
        s.set(sorMode=4, sorThreads=1024)
    b = watersim_hip.Batch(planets, graph=True, threads=1024)
    b.run(2)
    for s in ref:
        s.run_substeps(2)
    for s, n in zip(planets, sizes):
        assert s.get("residual") <= 1e-4
        p = s.download("px")
        assert np.all(np.isfinite(p)) and p.min() >= 0 and p.max() <= W
        assert np.isfinite(s.download("pvy")).all()
    for r, i in zip(ref, firsts):
        assert_same(r, planets[i], "planet %d (%d^3)" % (i, sizes[i]))
    print("  64 planets, %d particles; sweeps of the checked planets %s; %d planets XCD-local" % (
        sum(int(s.get("nP")) for s in planets), [int(planets[i].get("sweepsTotal")) for i in firsts],
        int(b.get("xcdLocalPlanets"))))
    b.close()
    for s in planets + ref:
        s.close()


def test_batch_refuses_timing_contexts():
    s = planet(32, 3)
    b = watersim_hip.Batch([s], graph=True, threads=256)
    s.set(timing=1)
    with pytest.raises(watersim_hip.WatersimError):
        b.run(1)
    s.set(timing=0)
    b.run(1)
    b.close()
    s.close()


def test_batch_with_balls_equals_single_planet():
    """Balls ride along in a batch: their raster, push-out and integrator run
    in the planets' own stages, and the planet ends where ws_run_substeps
    leaves it (balls included), eager and from the graph."""
    def with_balls(s, n):
        c = 0.5 * W
        r = 0.08 * W
        # x y z vx vy vz r rho submerged: two floating balls at the sea surface
        s.set_balls(np.array([[c + 0.4 * W, c, c, 0, 0, 0, r, 500.0, 0],
                              [c, c + 0.4 * W, c, 0.1, 0, 0, r, 700.0, 0]], np.float64))
        return s
    ref = with_balls(planet(32, 21), 32)
    ref.set(sorMode=4, sorThreads=256)
    bat = with_balls(planet(32, 21), 32)
    other = planet(48, 22)
    b = watersim_hip.Batch([bat, other], graph=True, threads=256)
    for k in (1, 3):
        b.run(k)
        ref.run_substeps(k)
        assert_same(ref, bat, "%d substeps" % k)
        assert np.array_equal(ref.balls(), bat.balls())
    print("  balls after 4 substeps: %s" % np.array2string(bat.balls()[:, :3], precision=4))
    b.close()
    for s in (ref, bat, other):
        s.close()
