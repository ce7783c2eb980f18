"""The Node host side: the N-API addon and the FluidSolver / MarchingTetrahedra
facade (watersim-v1_amd/js) run as plain `node` scripts, the way the
reference runs its own tests (test/*.js). The CPU check needs no GPU; the
others drive the HIP library through the addon on the GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JS = os.path.join(ROOT, "tests", "js")
ADDON = os.path.join(ROOT, "watersim-v1_amd", "_build", "watersim_napi.node")

pytestmark = pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")


def run_node(script, *args, timeout=300, gpu=False):
    env = dict(os.environ)
    if gpu:
        env["WATERSIM_EXPECT_GPU"] = "1"
    p = subprocess.run(["node", os.path.join(JS, script)] + list(args), cwd=ROOT, env=env, timeout=timeout,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    out = p.stdout.decode(errors="replace")
    print(out)
    assert p.returncode == 0, out[-3000:]
    return out


def test_addon_boundary_without_gpu():
    if not os.path.exists(ADDON):
        pytest.fail("the N-API addon is not built: run watersim-v1_amd/build.py")
    # where a GPU is visible (/dev/kfd) the no-GPU error path is not exercised
    run_node("boundary_cpu.test.js", timeout=120, gpu=os.path.exists("/dev/kfd"))


@pytest.mark.gpu
def test_marching_tets_through_node():
    run_node("surface_gpu.test.js", timeout=120, gpu=True)


@pytest.mark.gpu
def test_metaballs_through_node():
    run_node("metaball_gpu.test.js", timeout=120, gpu=True)


@pytest.mark.gpu
@pytest.mark.parametrize("section", ["settle", "heat", "shade", "currents", "whirl", "long", "balls"])
def test_planet_invariants_through_node(section):
    run_node("planet_gpu.test.js", section, timeout=110, gpu=True)


@pytest.mark.gpu
def test_optimization_invariants_through_node():
    """test/optimization.test.js:3-25 restated through the facade (heat
    conservation, the G2P cache against sampleVel, the gravity slider, the
    samplers' KAT), plus _markFluidCells, dryLandFraction, waveImpulse."""
    out = run_node("optimization_gpu.test.js", timeout=120, gpu=True)
    assert "ALL OPTIMIZATION INVARIANT TESTS PASSED" in out
