"""ws_tune (no reference counterpart): on a context's current system it
times the SOR kernels (the one-workgroup solve at three sizes, the
neighbour-synchronised solve), the P2G bins and, with the learned warm
start on, the network's two convolution tilings, keeps the fastest, and
leaves the simulation state as it found it (pressure, velocities,
particles and the solve counters)."""
import numpy as np
import pytest

from gpu_scenes import watersim_hip

pytestmark = pytest.mark.gpu

W = 5.1


def planet(n=32, seed=20260301):
    w = watersim_hip.build_world(n, W / n, 0.25 * W, 0.15 * W, seed, int(0.8 * n ** 3))
    s = watersim_hip.solver_for_world(w)
    s.set(pressureTol=1e-4, maxPressureIters=4000, residualEvery=8, vorticity=0.05)
    return s


def test_tune_keeps_state_and_picks_a_kernel():
    s = planet()
    s.run_substeps(3)
    s.sync()
    keys = ("q", "u", "v", "w", "px", "py", "pz", "pvx", "pvy", "pvz", "cellType")
    before = {k: s.download(k) for k in keys}
    counters = {k: s.get(k) for k in ("sweepsTotal", "residualsTotal", "pressureSolves", "sorMode", "warmStart")}
    rep = s.tune(16)
    print("  " + rep.replace("\n", "\n  "))
    modes = [ln for ln in rep.splitlines() if ln.startswith("sorMode ") and "sweeps in" in ln]
    assert len(modes) == 9  # modes 0-3, mode 4 at 256, 512 and 1024 threads, modes 5 and 6
    chosen, threads = int(s.get("sorMode")), int(s.get("sorThreads"))
    assert 0 <= chosen <= 6 and rep.rstrip().endswith("netTiled %d, p2gBin %d" % (int(s.get("netTiled")),
                                                                                  int(s.get("p2gBin"))))
    assert ("chosen: sorMode %d, sorThreads %d" % (chosen, threads)) in rep
    bins = {int(ln.split()[1]): float(ln.split(": ")[1].split()[0]) for ln in rep.splitlines()
            if ln.startswith("p2gBin ") and ln.rstrip().endswith(" ms")}
    assert sorted(bins) == [0, 1, 2, 3] and bins[int(s.get("p2gBin"))] == min(bins.values())
    times = {ln.split(":")[0]: float(ln.split(" in ")[1].split()[0]) for ln in modes}
    key = "sorMode %d (%d threads)" % (chosen, threads) if chosen in (4, 5) else "sorMode %d" % chosen
    assert times[key] == min(times.values())
    for k in keys:
        assert np.array_equal(s.download(k), before[k]), k
    for k, v in counters.items():
        if k != "sorMode":
            assert s.get(k) == v, k
    # the run goes on with the chosen kernel and still converges
    s.run_substeps(2)
    assert s.get("residual") <= 1e-4
    s.close()


def test_tune_rejects_bad_arguments():
    s = planet()
    with pytest.raises(watersim_hip.WatersimError):
        s.tune(0)
    s.close()
