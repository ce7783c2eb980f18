"""GPU frame tick against the reference on the thermal scene: heat
(exposure, relaxation, conduction, convection), evaporation with the
reference's _evS stream, currents and the phase pass, each stage from the
reference's input state. Flags, counts and stream states exact; float
fields within 1e-5 (FP32 kernels with expf/logf/cosf vs double)."""
import numpy as np
import pytest

import golden
from gpu_scenes import gpu_from_golden

pytestmark = pytest.mark.gpu


def load(tag):
    g = golden.load("thermal_32")
    s = gpu_from_golden(g, tag=tag)
    s.upload("cellType", g["t0_cellType"])
    s.upload("dens", g["t0_dens"])
    for k in "uvw":
        s.upload(k, g["t0_" + k])
    s.set_rng(g.meta["rng_before"]["cfS"], g.meta["rng_before"]["bfS"], g.meta["tick"]["evS_before"])
    return g, s


def close(s, g, tag, keys, tol=1e-5):
    for k in keys:
        a = s.download(k).astype(np.float64)
        b = g[tag + k].astype(np.float64)
        e = np.abs(a - b).max()
        print("  %-10s max |err| %.3e" % (tag + k, e))
        assert e <= tol, (tag + k, e)


def test_heat_tick():
    g, s = load("t0_")
    s.stage("heat", g.meta["tick"]["dtTick"])
    assert np.array_equal(s.download("pDepth"), g["t1_pDepth"])
    close(s, g, "t1_", ["pT", "pLight", "pvx", "pvy", "pvz"])
    s.close()


def test_evaporation_stream():
    g, s = load("t1_")
    s.stage("evaporation", g.meta["tick"]["dtTick"])
    print("  evaporated gpu %d ref %d" % (s.get("evapCount"), g.meta["tick"]["evapCount"]))
    assert int(s.get("evapCount")) == g.meta["tick"]["evapCount"]
    assert int(s.get_rng()[2]) == g.meta["tick"]["evS_after"]
    assert np.array_equal(s.download("pflag"), g["t2_pflag"])
    close(s, g, "t2_", ["pvx", "pvy", "pvz", "pAir"])
    s.close()


def test_currents():
    g, s = load("t2_")
    s.stage("currents", g.meta["tick"]["dtTick"])
    close(s, g, "t3_", ["pvx", "pvy", "pvz"])
    s.close()


def test_phase_changes():
    g, s = load("t3_")
    s.stage("phase", 1 / 60)
    assert np.array_equal(s.download("pflag"), g["t4_pflag"])
    close(s, g, "t4_", ["pvx", "pvy", "pvz", "pAir", "pWx", "pWy", "pWz", "pPh"])
    s.close()


def test_push_surface():
    """_pushSurfaceParticles on the perturbed frame-end state: every rescued
    particle lands where the reference puts it (positions are walked in
    float here, in double there; the rock tests agree, so the results match
    to float rounding)."""
    g = golden.load("thermal_32")
    s = gpu_from_golden(g, tag="fe2_")
    s.stage("push_surface", 0.0)
    moved = 0
    for k in ["px", "py", "pz", "pvx", "pvy", "pvz"]:
        a = s.download(k).astype(np.float64)
        b = g["fe3_" + k].astype(np.float64)
        e = np.abs(a - b).max() / max(1.0, np.abs(b).max())
        moved += int(np.sum(g["fe2_" + k] != g["fe3_" + k]))
        print("  %-6s max rel err %.3e" % (k, e))
        assert e <= 2e-5, (k, e)
    assert moved > 100
    s.close()


def _fe_gpu(tag):
    g = golden.load("thermal_32")
    s = gpu_from_golden(g, tag=tag)
    fe = g.meta["frame_end"]
    sp = fe["sunPos"]
    s.set(spinOn=1, hasSun=1, sunX=sp[0], sunY=sp[1], sunZ=sp[2])
    return g, s


def _fe_compare(s, g, a, b, keys, tol=2e-5):
    moved = 0
    for k in keys:
        got = s.download(k).astype(np.float64)
        want = g[b + k].astype(np.float64)
        e = np.abs(got - want).max() / max(1.0, np.abs(want).max())
        moved += int(np.sum(g[a + k] != g[b + k]))
        print("  %-5s max rel err %.3e" % (k, e))
        assert e <= tol, (k, e)
    return moved


@pytest.mark.parametrize("serial_max,queue_max,lds_max", [(48, 16000, 16384), (0, 16000, 16384),
                                                          (1 << 30, 16000, 16384), (0, 0, 16384), (0, 0, 0)])
def test_vapor_collisions(serial_max, queue_max, lds_max):
    """Parcel-parcel and parcel-liquid contacts on the perturbed frame-end
    state (a chain of 80 touching parcels along the sea surface, real steam
    above it): the census exact, positions, velocities and spins to float
    rounding of the reference's serial sweep. The settings pick the path of
    the contact groups: serial_max 0 sends every group to the
    dependency-ordered sweep, 2^30 every group to the one-thread serial
    sweep; queue_max 0 moves the dependency-ordered groups from the LDS ready
    queue to the polling sweep, whose counts lds_max 0 moves from LDS to
    global memory."""
    g, s = _fe_gpu("fe0_")
    s.set(vaporSerialMax=serial_max, vaporQueueMax=queue_max, vaporLdsMax=lds_max)
    s.stage("vapor_collisions", 0.0)
    fe = g.meta["frame_end"]
    assert int(s.get("vaporCount")) == fe["vaporCount"]
    assert int(s.get("nightVaporCount")) == fe["nightVaporCount"]
    assert s.get("contactOverflow") == 0
    moved = _fe_compare(s, g, "fe0_", "fe1_", ["px", "py", "pz", "pvx", "pvy", "pvz", "pWx", "pWy", "pWz"])
    assert moved > 100
    s.close()


def test_ice_contacts():
    g, s = _fe_gpu("fe1_")
    s.stage("ice_contacts", 0.0)
    assert np.array_equal(s.download("pAir"), g["fe2_pAir"])
    moved = _fe_compare(s, g, "fe1_", "fe2_", ["pvx", "pvy", "pvz"])
    assert moved > 20
    s.close()


def test_vapor_liquid_crowded_parcel():
    """A parcel with 200 liquid particles between 1.5 and 3 contact radii
    and three inside one radius (more candidates than the 64 the first
    version of the pass kept): the GPU pass meets them in index order as the
    oracle's sweep does, raises no overflow, and gives the same bits on a
    second context."""
    import cpu_oracle  # noqa: F401  (the oracle library must be built)
    from scenes import oracle_from_golden, PART_KEYS, EXTRA_KEYS
    g = golden.load("thermal_32")
    fe = g.meta["frame_end"]
    sp = fe["sunPos"]
    w = g.meta["world"]
    rad = 0.6 * w["spacing"]
    n = len(g["fe0_px"])
    keys = [k for k in PART_KEYS + EXTRA_KEYS if "fe0_" + k in g]
    parts = {k: g["fe0_" + k].copy() for k in keys}
    fl = parts["pflag"]
    vap = np.nonzero(fl == 2)[0]
    v = int(vap[len(vap) // 2])
    rng = np.random.default_rng(7)
    K = 203
    d = rng.normal(size=(K, 3))
    r = np.concatenate([rng.uniform(1.5, 3.0, size=200), [0.5, 0.7, 0.9]]) * rad
    d *= (r / np.linalg.norm(d, axis=1))[:, None]
    liq = int(np.nonzero(fl == 0)[0][0])
    for k in keys:
        extra = np.repeat(parts[k][liq:liq + 1], K)
        if k in ("px", "py", "pz"):
            extra = (parts[k][v] + d[:, "xyz".index(k[1])]).astype(parts[k].dtype)
        parts[k] = np.concatenate([parts[k], extra])
    N = n + K
    outs = []
    for rep in range(2):
        s = gpu_from_golden(g, tag="fe0_", capacity=N)
        s.set(nP=N, spinOn=1, hasSun=1, sunX=sp[0], sunY=sp[1], sunZ=sp[2])
        for k in keys:
            s.upload(k, parts[k])
        s.stage("vapor_collisions", 0.0)
        assert s.get("contactOverflow") == 0
        outs.append({k: s.download(k) for k in ("px", "py", "pz", "pvx", "pvy", "pvz")})
        s.close()
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
    st = oracle_from_golden(g, tag="fe0_")
    st.set_particles(N, **parts)
    st.set_params(spinOn=1, hasSun=1, sunX=sp[0], sunY=sp[1], sunZ=sp[2])
    st.call("oc_vapor_collisions")
    moved = 0
    for k in outs[0]:
        want = st[k][:N].astype(np.float64)
        e = np.abs(outs[0][k] - want).max() / max(1.0, np.abs(want).max())
        print("  %-4s max rel err vs oracle %.3e" % (k, e))
        assert e <= 2e-5, (k, e)
        moved += int(abs(float(outs[0][k][v]) - float(parts[k][v])) > 0)
    assert moved > 0  # the crowded parcel was pushed


def test_vapor_liquid_chain_walks_a_parcel():
    """A chain of eight spray particles along y, in index order, each 0.7
    contact radii beyond the previous one: every contact pushes the parcel
    (moved to the empty centre of the planet) into reach of the next, 5.4
    radii in all, more than the three radii gathered around its starting
    point. The pass gathers again around the parcel once it has moved two
    radii, so it meets the whole chain as the oracle's sweep does (the chain
    runs across the x window of the reference's sorted vapor list, which
    therefore stays valid)."""
    from scenes import oracle_from_golden, PART_KEYS, EXTRA_KEYS
    g = golden.load("thermal_32")
    fe = g.meta["frame_end"]
    sp = fe["sunPos"]
    w = g.meta["world"]
    rad = 0.6 * w["spacing"]
    c = 0.5 * w["nx"] * w["dx"]
    n = len(g["fe0_px"])
    keys = [k for k in PART_KEYS + EXTRA_KEYS if "fe0_" + k in g]
    parts = {k: g["fe0_" + k].copy() for k in keys}
    fl = parts["pflag"]
    v = int(np.nonzero(fl == 2)[0][0])
    parts["px"][v], parts["py"][v], parts["pz"][v] = c, c, c
    parts["pvx"][v] = parts["pvy"][v] = parts["pvz"][v] = 0.0
    K = 8
    liq = int(np.nonzero(fl == 0)[0][0])
    for k in keys:
        extra = np.repeat(parts[k][liq:liq + 1], K)
        if k == "px" or k == "pz":
            extra[:] = c
        elif k == "py":
            extra = (c + (-0.5 + 0.7 * np.arange(K)) * rad).astype(parts[k].dtype)
        elif k == "pflag":
            extra[:] = 1
        elif k in ("pvx", "pvy", "pvz"):
            extra[:] = 0.0
        parts[k] = np.concatenate([parts[k], extra])
    N = n + K
    s = gpu_from_golden(g, tag="fe0_", capacity=N)
    s.set(nP=N, spinOn=1, hasSun=1, sunX=sp[0], sunY=sp[1], sunZ=sp[2])
    for k in keys:
        s.upload(k, parts[k])
    s.stage("vapor_collisions", 0.0)
    assert s.get("contactOverflow") == 0
    got = {k: s.download(k) for k in ("px", "py", "pz", "pvx", "pvy", "pvz")}
    st = oracle_from_golden(g, tag="fe0_")
    st.set_particles(N, **parts)
    st.set_params(spinOn=1, hasSun=1, sunX=sp[0], sunY=sp[1], sunZ=sp[2])
    st.call("oc_vapor_collisions")
    walk = (float(st["py"][v]) - c) / rad
    print("  parcel walked %.3f radii (oracle), %.3f (gpu)" % (walk, (float(got["py"][v]) - c) / rad))
    assert walk > 5.0
    for k in got:
        want = st[k][:N].astype(np.float64)
        e = np.abs(got[k] - want).max() / max(1.0, np.abs(want).max())
        print("  %-4s max rel err vs oracle %.3e" % (k, e))
        assert e <= 2e-5, (k, e)


LCG_A, LCG_C = 1664525, 1013904223


def lcg_back(s, n):
    """The stream state n draws before s."""
    inv = pow(LCG_A, -1, 1 << 32)
    for _ in range(n):
        s = ((s - LCG_C) * inv) % (1 << 32)
    return s


@pytest.mark.parametrize("dup,rate,reject_at", [(3, 0.02, None), (3, 1.0, None), (1, 0.3, None), (3, 1e9, None),
                                                 (1, 1e9, 19), (10, 1.0, None), (10, 1e9, None)])
def test_evaporation_windows(dup, rate, reject_at):
    """The chooser tests a window of candidates under several counts of
    earlier ejections at once; this drives each way a window ends against
    the oracle's serial walk of the _evS stream: the scene's particles
    repeated dup times (5151 candidates, more than one window of 4096), the
    scene's rate times `rate` (0.02: 14 ejections, so whole windows; 1.0 and
    0.3: about one candidate in ten and one in thirty; 1e9: every candidate
    ejects, so windows end after their sixteenth ejection), and a start
    state whose stream puts a value below
    the 1e-9 cut at the first normal of candidate reject_at (every candidate
    ejecting takes seven draws, that one eight). With dup 10 (17 170
    candidates) a run of candidates crosses the chooser's 16 384-threshold
    staging in LDS. Counts, the stream state after the pass and the flags
    exact; velocities within 1e-5."""
    import cpu_oracle  # noqa: F401  (the oracle library must be built)
    from scenes import oracle_from_golden, tick_params, PART_KEYS, EXTRA_KEYS
    g = golden.load("thermal_32")
    keys = [k for k in PART_KEYS + EXTRA_KEYS if "t1_" + k in g]
    parts = {k: np.concatenate([g["t1_" + k]] * dup) for k in keys}
    N = len(parts["px"])
    act = g.meta["tick"]["sunActivity"] * rate
    ev0 = g.meta["tick"]["evS_before"] if reject_at is None else lcg_back(2, 7 * reject_at + 2)
    dt = g.meta["tick"]["dtTick"]
    s = gpu_from_golden(g, tag="t1_", capacity=N)
    s.set(nP=N)
    for k in keys:
        s.upload(k, parts[k])
    s.upload("cellType", g["t0_cellType"])
    s.upload("dens", g["t0_dens"])
    for k in "uvw":
        s.upload(k, g["t0_" + k])
    rb = g.meta["rng_before"]
    s.set_rng(rb["cfS"], rb["bfS"], ev0)
    s.set(sunActivity=act)
    s.stage("evaporation", dt)
    st = oracle_from_golden(g, tag="t1_")
    st.set_particles(N, **parts)
    st.arrays["cellType"][:] = g["t0_cellType"]
    st.arrays["dens"][:] = g["t0_dens"]
    for k in "uvw":
        st.arrays[k][:] = g["t0_" + k]
    st.set_params(**tick_params(g))
    st.set_params(sunActivity=act, evS=ev0)
    st.call("oc_update_evaporation", dt)
    n_gpu, n_ref = int(s.get("evapCount")), st.get_i("evapCount")
    print("  %d particles: evaporated gpu %d oracle %d" % (N, n_gpu, n_ref))
    assert n_gpu == n_ref and n_ref > 0
    assert int(s.get_rng()[2]) == st.get_i("evS")
    assert np.array_equal(s.download("pflag"), st["pflag"][:N])
    for k in ("pvx", "pvy", "pvz", "pAir"):
        e = np.abs(s.download(k).astype(np.float64) - st[k][:N].astype(np.float64)).max()
        print("  %-4s max |err| %.3e" % (k, e))
        assert e <= 1e-5, (k, e)
    s.close()


@pytest.mark.parametrize("offset", [0, 1])
def test_evaporation_draw_at_threshold(offset):
    """A candidate whose draw is exactly the largest stream state that ejects
    (state / 2^32 == floor(prob 2^32) / 2^32 <= prob: ejects) and one state
    above it (does not): the probability and its threshold are evaluated in
    double as _evapChunk does (solver.js:3062-3067). One particle the
    reference ejected in this tick is the only candidate (every other one is
    taken out of the pass by its flag); GPU and oracle agree on the flag,
    the count and the stream state after the pass."""
    import math
    from scenes import oracle_from_golden, tick_params, PART_KEYS, EXTRA_KEYS
    g = golden.load("thermal_32")
    keys = [k for k in PART_KEYS + EXTRA_KEYS if "t1_" + k in g]
    ejected = np.flatnonzero((g["t1_pflag"] == 0) & (g["t2_pflag"] == 2))
    p = int(ejected[0])
    parts = {k: g["t1_" + k].copy() for k in keys}
    parts["pflag"][:] = 1
    parts["pflag"][p] = 0
    t = g.meta["tick"]
    dt = t["dtTick"]
    rate = t["sunActivity"] * t["evapIntensity"]
    freeze = max(t["meltT"], t["snowT"] * 1.05)
    prob = min(1.0, rate * dt * math.exp(3.5 * (float(parts["pT"][p]) - freeze)))
    lim = math.floor(prob * 4294967296.0)
    ev0 = lcg_back(lim + offset, 1)
    N = len(parts["px"])
    s = gpu_from_golden(g, tag="t1_")
    for k in keys:
        s.upload(k, parts[k])
    s.upload("cellType", g["t0_cellType"])
    for k in "uvw":
        s.upload(k, g["t0_" + k])
    rb = g.meta["rng_before"]
    s.set_rng(rb["cfS"], rb["bfS"], ev0)
    s.stage("evaporation", dt)
    st = oracle_from_golden(g, tag="t1_")
    st.set_particles(N, **parts)
    st.arrays["cellType"][:] = g["t0_cellType"]
    for k in "uvw":
        st.arrays[k][:] = g["t0_" + k]
    st.set_params(**tick_params(g))
    st.set_params(evS=ev0)
    st.call("oc_update_evaporation", dt)
    want = 1 if offset == 0 else 0
    print("  particle %d, prob %.17g, threshold %d, draw threshold + %d: gpu %d oracle %d"
          % (p, prob, lim, offset, s.get("evapCount"), st.get_i("evapCount")))
    assert st.get_i("evapCount") == want
    assert int(s.get("evapCount")) == want
    assert int(s.download("pflag")[p]) == int(st["pflag"][p]) == (2 if want else 0)
    assert int(s.get_rng()[2]) == st.get_i("evS")
    s.close()
