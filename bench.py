#!/usr/bin/env python3
"""Benchmark: FLIP substeps/s of the planet substep pipeline on one MI355X.

Workload (default "cfg3", full physics): 128^3 MAC grid, ~2.0M FLIP
particles, seeded voxel planet (core 0.25 W, sea level 0.4 W, W = 5.1 m),
FLIP 0.95, red-black SOR (omega 1.7) to relative residual 1e-4, vorticity
confinement 0.05, dt from the CFL law, driven frame by frame (1/60 s frames)
as FluidSolver.step does: the 25 Hz thermal / evaporation tick, the substep
loop and the frame-end phase pass. Particle temperatures start from a
latitude profile of 270-300 K; the sun circles the planet once per 2000
substeps at 23 degrees declination. Synthetic input built on the host by the
library's world builder; all state is resident in HBM before timing.

A "step" is one substep. Prints one JSON line (diagnostics on stderr).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "watersim-v1_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

W_DOMAIN = 5.1
CONFIGS = {
    # name: grid n, particle target, seed, vorticity, SOR mode, frame-driven full physics
    "cfg1": dict(n=32, target=10000, seed=20260114, vorticity=0.0, tol=0.0, iters=40, physics=False, tune=True),
    "cfg2": dict(n=64, target=232000, seed=20260115, vorticity=0.05, tol=1e-4, iters=0, physics=False,
                 frames=True, currents=0.8, curl_init=0.5, surface_every=10, tune=True),
    "cfg3": dict(n=128, target=2117000, seed=20260116, vorticity=0.05, tol=1e-4, iters=0, physics=True, tune=True),
    # the learned warm start on the cfg3 scene (a seed held out of training)
    "cfg4": dict(n=128, target=2117000, seed=20260116, vorticity=0.05, tol=1e-4, iters=0, physics=True,
                 learned=True),
    # 64 independent planets of 32^3..96^3, about 8 particles per liquid cell
    "cfg5": dict(planets=64, sizes=(32, 48, 64, 96), per_cell=0.8, seed=20260300, vorticity=0.05, tol=1e-4,
                 batched=True),
}
WEIGHTS = os.path.join(ROOT, "watersim-v1_amd", "weights", "warmstart_unet.f32")
BF16_PEAK = 2.5e15  # MI355X dense BF16 matrix FLOP/s


def net_flops(n, coarse=False):
    """Multiply-adds x 2 of one inference of the warm-start network on an
    n^3 solver grid (the network's grid m = n, or n/2 when it runs on the
    coarsened grid): 3x3x3 convolutions 5->16->16 (m^3), 16->32->32
    ((m/2)^3), 32->64->64 ((m/4)^3), 96->32->32 ((m/2)^3), 48->16->16 (m^3),
    head."""
    V = (n // 2 if coarse else n) ** 3
    full = 16 * 5 + 16 * 16 + 48 * 16 + 16 * 16
    half = 16 * 32 + 32 * 32 + 96 * 32 + 32 * 32
    quarter = 32 * 64 + 64 * 64
    return 2 * 27 * (V * full + V / 8 * half + V / 64 * quarter) + 2 * 16 * V


def net_bytes(n, coarse=False):
    """Algorithmic HBM bytes of one inference on an n^3 solver grid: every
    layer reads its bf16 input once and writes its bf16 output once (the 27
    taps reuse the input through LDS), pooling and up-sampling read and write
    their tensors once. Per network voxel (grid m = n, or n/2 coarsened):
    500 B at m^3 (convolutions 0, 1, 8, 9, head, pooling read, skip
    concatenation), 1024 B at (m/2)^3, 640 B at (m/4)^3. The solver side at
    full resolution: 56 B per cell (residual, input tensor, update), or on
    the coarsened grid 32 B per cell (residual; the block averages read the
    cell types, r0 and q; the update reads the types and updates q) and the
    16-byte input voxel."""
    V = n ** 3
    m = (n // 2 if coarse else n) ** 3
    side = V * 32 + m * 16 if coarse else V * 56
    return m * 500 + m / 8 * 1024 + m / 64 * 640 + side


FRAME = 1.0 / 60
SUN_PERIOD = 2000      # substeps per revolution of the sun
SUN_DECL = 23.0        # degrees between the sun's path and the planet's equator
SUN_ACTIVITY = 0.6     # js/solver.js:142 "default daylight"


def kelvin_to_model(tk):
    """Reference temperatures are normalised (0 .. 1.15, ambient 0.32, snow
    0.10, melt 0.105; js/solver.js:158-190). 273.15 K maps to the melt point
    and 60 K span one unit: T = 0.105 + (T_K - 273.15) / 60."""
    return 0.105 + (tk - 273.15) / 60.0
HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_solver(cfg):
    import watersim_hip
    n = cfg["n"]
    dx = W_DOMAIN / n
    t = time.time()
    world = watersim_hip.build_world(n, dx, 0.25 * W_DOMAIN, 0.15 * W_DOMAIN, cfg["seed"], cfg["target"])
    log("world %d^3: %d particles, land %.3f, built in %.1fs" % (n, world["nP"], world["landFrac"], time.time() - t))
    s = watersim_hip.solver_for_world(world)
    s.set(pic=0.05, omega=1.7, gravity=9.8, maxSpeed=3.2, vorticity=cfg["vorticity"], cfl=3.0)
    if cfg["physics"]:
        # latitude profile 270 K (poles) .. 300 K (equator) about the y axis
        c = 0.5 * W_DOMAIN
        x, y, z = (s.download(k)[:world["nP"]].astype(np.float64) - c for k in ("px", "py", "pz"))
        r = np.sqrt(x * x + y * y + z * z)
        sin_lat = np.where(r > 0, y / np.maximum(r, 1e-12), 0.0)
        s.upload("pT", kelvin_to_model(270.0 + 30.0 * (1.0 - sin_lat ** 2)).astype(np.float32))
        s.set(sunActivity=SUN_ACTIVITY, hasSun=1)
        set_sun(s, world, 0)
    if cfg.get("currents"):
        s.set(currents=cfg["currents"])
    if cfg.get("curl_init"):
        nP = world["nP"]
        x, y, z = (s.download(k)[:nP].astype(np.float64) for k in ("px", "py", "pz"))
        u = curl_noise_velocity(x, y, z, cfg["seed"], cfg["curl_init"], W_DOMAIN)
        for c, k in enumerate(("pvx", "pvy", "pvz")):
            s.upload(k, u[c])
    if cfg["tol"] > 0:
        s.set(pressureTol=cfg["tol"], maxPressureIters=4000, residualEvery=8)
    else:
        s.set(pressureIters=cfg["iters"])
    return s, world


def curl_noise_velocity(x, y, z, seed, rms, width, modes=24):
    """Divergence-free random velocity: the analytic curl of a seeded sum of
    plane-wave vector potentials, psi = sum_m a_m sin(k_m . x + phi_m), so
    u = sum_m cos(k_m . x + phi_m) (k_m x a_m) has zero divergence exactly.
    Wavelengths span width/4 .. width/1.5; scaled to the given RMS speed."""
    rng = np.random.default_rng(seed)
    u = np.zeros((3, len(x)))
    for _ in range(modes):
        d = rng.normal(size=3)
        k = d / np.linalg.norm(d) * 2 * np.pi / width * rng.uniform(1.5, 4.0)
        a = rng.normal(size=3)
        th = k[0] * x + k[1] * y + k[2] * z + rng.uniform(0, 2 * np.pi)
        u += np.cross(k, a)[:, None] * np.cos(th)[None, :]
    u *= rms / max(np.sqrt((u * u).sum(0).mean()), 1e-30)
    return u.astype(np.float32)


def set_sun(s, world, substeps_done):
    """Sun position in the planet frame: distance 3.4 oceanR from the centre
    (js/scene.js:1248), hour angle advancing one turn per SUN_PERIOD substeps."""
    c = 0.5 * W_DOMAIN
    th = 2 * np.pi * substeps_done / SUN_PERIOD
    de = np.radians(SUN_DECL)
    d = 3.4 * world["oceanR"]
    s.set(sunX=c + d * np.cos(de) * np.cos(th), sunY=c + d * np.sin(de), sunZ=c + d * np.cos(de) * np.sin(th))


def advance(s, world, cfg, n, done):
    """n substeps: frames through the full step (tick, substeps, phase pass)
    when the workload has the thermal cycle, else back-to-back substeps with
    the CFL law evaluated on the device. Returns the substeps executed."""
    if not (cfg["physics"] or cfg.get("frames")):
        s.run_substeps(n)
        return n
    k, every = 0, cfg.get("surface_every", 0)
    while k < n:
        if cfg["physics"]:
            set_sun(s, world, done + k)
        m = s.step(FRAME)
        if (k + m) // 250 != k // 250:
            log("  %d substeps" % (k + m))
        if every and (done + k) // every != (done + k + m) // every:
            # the surface, iso calibrated to the water volume (js/solver.js:3838-3849)
            s.splat_density()
            s.mt_build(iso=s.autotune_iso())
        k += m
    return k


def algorithmic_bytes(cfg, n, nP, cell_sweeps, residuals, ticks):
    """Bytes the reference dataflow moves (each array once per stage at its
    reference precision): plain substep 154 N + 69 F + 16 C, 13 per cell per
    SOR sweep (the reference sweeps its fluid list, solver.js:1397-1404),
    9 C per residual evaluation, vorticity 12 F + 82 C, and per thermal tick
    115 N + 77 C + 4 D plus evaporation 17 N + C."""
    C, F, D = n ** 3, 3 * n * n * (n + 1), (n + 1) ** 3
    b = 154 * nP + 69 * F + 16 * C
    if cfg["vorticity"] > 0:
        b += 12 * F + 82 * C
    return b, 13 * cell_sweeps + 9 * C * residuals + ticks * (132 * nP + 78 * C + 4 * D)


def cpu_baseline(s, world, cfg, budget_s=15.0, max_substeps=8):
    """The oracle (single-threaded C port of the reference: reference
    arithmetic, lexicographic SOR) on the same state, for a bounded sample.
    With the thermal cycle on it runs whole frames as js/solver.js:2670-2700
    does: the tick when 1/25 s has accumulated, the substeps, the phase pass."""
    import cpu_oracle as oracle
    n = cfg["n"]
    st = oracle.OracleState({"nx": n, "ny": n, "nz": n, "dx": W_DOMAIN / n, "coreR": world["coreR"],
                             "oceanR": world["oceanR"], "spacing": world["spacing"],
                             "particleVolume": world["particleVolume"], "dv": world["dv"], "tn": world["tn"]})
    st.set_static(world["solid"], world["R"], world["staticTypes"], world["gu"], world["gv"], world["gw"])
    nP = int(s.get("nP"))
    parts = {k: s.download(k) for k in ["px", "py", "pz", "pvx", "pvy", "pvz", "pT", "pAir", "pflag", "pcool",
                                        "pLight", "pDepth", "pWx", "pWy", "pWz", "pPh"]}
    st.set_particles(nP, **parts)
    st.arrays["q"][:] = s.download("q")
    st.set_params(pic=0.05, omega=1.7, gravity=9.8, maxSpeed=3.2, vorticity=cfg["vorticity"], cfl=3.0,
                  gRamp=s.get("gRamp"), umax=s.get("umax"))
    if cfg["tol"] > 0:
        st.set_params(pressureTol=cfg["tol"], maxPressureIters=4000, residualEvery=8, sorRedBlack=0)
    else:
        st.set_params(pressureIters=cfg["iters"])
    done, t0 = 0, time.perf_counter()
    if cfg["physics"]:
        st.set_params(hasSun=1, sunActivity=SUN_ACTIVITY, Rlo=world["Rlo"], Rhi=world["Rhi"])
        sim, acc = s.get("simTime"), s.get("heatAcc")
        while done < max_substeps and time.perf_counter() - t0 < budget_s:
            c = 0.5 * W_DOMAIN
            th, de, d = 2 * np.pi * done / SUN_PERIOD, np.radians(SUN_DECL), 3.4 * world["oceanR"]
            st.set_params(sunX=c + d * np.cos(de) * np.cos(th), sunY=c + d * np.sin(de),
                          sunZ=c + d * np.cos(de) * np.sin(th))
            sim += FRAME
            acc += FRAME
            st.set_params(simTime=sim)
            if acc >= 1.0 / 25:
                tick = min(acc, 1.0 / 12)
                st.call("oc_update_heat", ctypes.c_double(tick))
                st.call("oc_update_evaporation", ctypes.c_double(tick))
                acc = 0.0
            rem, k = FRAME, 0
            while rem > 1e-6 and k < 10:
                dt = min(FRAME, st.call("oc_dt_law"))
                st.call("oc_substep", ctypes.c_double(dt))
                rem -= dt
                k += 1
            done += k
            st.call("oc_update_phase", ctypes.c_double(FRAME))
        what = "frames (tick, substeps, phase pass)"
    else:
        while done < max_substeps and time.perf_counter() - t0 < budget_s:
            st.call("oc_substep", st.call("oc_dt_law"))
            done += 1
        what = "substeps"
    el = time.perf_counter() - t0
    out = {"value": done / el, "unit": "substeps/s", "cores": 1, "kind": "port",
           "sample": "%d substeps run as %s from the same %d^3 / %d-particle state (oracle: reference "
                     "arithmetic, lexicographic SOR %s), single thread"
                     % (done, what, n, nP, "to residual %g" % cfg["tol"] if cfg["tol"] else "%d sweeps" % cfg["iters"]),
           "cpu": cpu_model()}
    # the reference's own single-threaded path (Node) against this port on
    # identical workloads, both timed in the development container
    # (tests/golden/time_reference.py): the port's time there over the
    # reference's, so the reference on this host is about value * ratio
    rt = os.path.join(ROOT, "tests", "golden", "reference_timing.json")
    if os.path.exists(rt):
        with open(rt) as f:
            ref = json.load(f)
        ratios = {k: v["twin_over_reference"] for k, v in ref["configs"].items()}
        ratio = sum(ratios.values()) / len(ratios)
        out["reference_equiv_value"] = out["value"] * ratio
        out["port_over_reference_time"] = ratios
        out["reference_timed_on"] = "%s, Node %s, 1 thread: %s" % (ref["cpu"], ref["node"], ", ".join(
            "%s %.2f ms/substep (port %.2f)" % (k, v["reference"]["median_ms"], v["twin"]["median_ms"])
            for k, v in ref["configs"].items()))
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed substeps")
    ap.add_argument("--warmup", type=int, default=20, help="untimed substeps")
    ap.add_argument("--config", default="cfg3", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain solve pass (cfg3)")
    ap.add_argument("--weights", default=WEIGHTS, help="warm-start network weights (cfg4)")
    ap.add_argument("--sor-mode", type=int, default=None,
                    help="use this SOR kernel instead of the autotuner's choice (a profiler changes the timings the "
                         "autotuner compares)")
    ap.add_argument("--p2g-bin", type=int, default=None, help="use this P2G bin instead of the autotuner's choice")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB five times and exit: the "
                         "reference kernel whose FETCH_SIZE / WRITE_SIZE counters scale the HBM traffic")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch
        nbytes = a.copy_calibration << 20
        src = torch.ones(nbytes // 4, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 5}), flush=True)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: the planet substep runs on a single MI355X; "
              "--gpus %d is not supported (a planet would shard by k-slabs, batched planets by planet)"
              % a.gpus, file=sys.stderr)
        sys.exit(2)
    cfg = CONFIGS[a.config]
    if cfg.get("learned"):
        return learned_main(a, cfg)
    if cfg.get("batched"):
        return batched_main(a, cfg)
    s, world = make_solver(cfg)
    n, nP = cfg["n"], int(s.get("nP"))
    log("warmup %d substeps" % a.warmup)
    done = advance(s, world, cfg, a.warmup, 0)
    s.sync()
    tune_report = None
    if cfg.get("tune"):
        # the autotuner picks the SOR kernel and the P2G bin for this grid
        # (outside the timed region)
        tune_report = s.tune(64)
        log("tune: %s" % tune_report.replace("\n", "; "))
    if a.sor_mode is not None:
        s.set(sorMode=a.sor_mode)
        tune_report = (tune_report + "; " if tune_report else "") + "sorMode %d set by --sor-mode" % a.sor_mode
    if a.p2g_bin is not None:
        s.set(p2gBin=a.p2g_bin)
        tune_report = (tune_report + "; " if tune_report else "") + "p2gBin %d set by --p2g-bin" % a.p2g_bin
    # HIP events around the solver's SOR batches and pressure solves stay
    # queued on its stream and are read after the timed region (no added
    # host synchronization)
    s.set(timing=1)
    log("timing %d substeps" % a.steps)
    s.sync()
    t0 = time.perf_counter()
    steps = advance(s, world, cfg, a.steps, done)
    s.sync()
    el = time.perf_counter() - t0
    g = {k: s.get(k) for k in ("sorMs", "sorLaunches", "pressureMs", "pressureSolves", "sweepsTotal",
                               "residualsTotal", "ticksTotal", "sorMode", "sorCellSweeps")}
    s.set(timing=0)
    sps = steps / el
    C = n ** 3
    # dominant kernel: the SOR sweep. Algorithmic bytes: 13 per cell updated
    # (q read + write, rhs, cell type). The list kernel updates the fluid
    # cells of one colour per launch, the tiled kernel every cell of a full
    # sweep, the grid kernel every cell of one colour.
    mode = int(g["sorMode"])
    launches = max(g["sorLaunches"], 1)
    sor_avg_s = (g["sorMs"] / 1e3) / launches
    cell_sweeps = g["sorCellSweeps"]
    if mode in (2, 4, 5, 6):
        sor_bytes = 13.0 * g["sorCellSweeps"] / launches
    else:
        # the kernels over the whole grid count every cell; the algorithmic
        # work is the fluid cells (the reference sweeps its fluid list):
        # 13 B per fluid cell and sweep, fluid cells from the final state
        nfl = float(np.sum(s.download("cellType") == 1))
        per_launch = {0: 0.5, 1: 1.0, 3: 2.0}[mode] * nfl
        sor_bytes = 13.0 * per_launch
        cell_sweeps = g["sweepsTotal"] * nfl
    achieved = sor_bytes / sor_avg_s / 1e9 if g["sorLaunches"] else 0.0
    kname = {6: "k_sor_xcd", 5: "k_sor_flow", 4: "k_sor_solve", 3: "k_sor_tb", 2: "k_sor_list", 1: "k_sor_tile",
             0: "k_sor"}[mode]
    kdesc = {6: " (the red-black sweeps between two residual checks, one launch on the workgroups of one XCD)",
             5: " (a whole pressure solve on workgroups that wait only for their neighbours)",
             4: " (a whole pressure solve in one workgroup: its sweeps and residual checks)",
             3: " (two red-black sweeps of the fluid cells per launch, LDS tiles)",
             2: " (the fluid cells of one colour)", 1: " (one full red-black sweep, LDS tile)",
             0: " (one colour of the whole grid; algorithmic bytes count its fluid cells)"}[mode]
    traffic = read_traffic(kname, a.config)
    b_plain, b_iter = algorithmic_bytes(cfg, n, nP, cell_sweeps, g["residualsTotal"], g["ticksTotal"])
    b_total = b_plain * steps + b_iter
    ke = s.kinetic_energy()
    # outside the timed region: one render readback of the fields a renderer
    # reads each frame (PCIe-inclusive; never part of `value`)
    s.read_render()
    t1 = time.perf_counter()
    for _ in range(3):
        s.read_render()
    readback_ms = 1e3 * (time.perf_counter() - t1) / 3
    log("substep %.3f ms, pressure %.3f ms/solve, %.1f sweeps/substep, %d ticks, SOR launch %.2f us, ke %.4g"
        % (1e3 * el / steps, g["pressureMs"] / max(g["pressureSolves"], 1), g["sweepsTotal"] / steps,
           g["ticksTotal"], 1e6 * sor_avg_s, ke))
    learned = None
    if a.config == "cfg3" and os.path.exists(a.weights) and not a.no_learned:
        # outside the timed region: the metric's second half, pressure-solve
        # ms to residual 1e-4 with the learned warm start against plain SOR,
        # both from the same assembled system on the following substeps
        learned = learned_compare(s, world, cfg, a.weights, 6, done + steps)
    sor_txt = "to residual %g" % cfg["tol"] if cfg["tol"] else "%d sweeps" % cfg["iters"]
    window = "timed substeps %d-%d of a run from the fill at rest" % (done + 1, done + steps)
    if cfg["physics"]:
        wl = ("%s: %d^3 MAC grid, %d FLIP particles, red-black SOR %s, vorticity %g, full thermal cycle "
              "(25 Hz heat + evaporation tick, frame-end vapor collisions, phase pass, snow contacts and surface "
              "rescue, rotating sun, 270-300 K latitude profile), "
              "1/60 s frames; %s (the vapor layer grows over the 1000-substep run and the frame-end contact "
              "passes with it)" % (a.config, n, nP, sor_txt, cfg["vorticity"], window))
    elif cfg.get("frames"):
        wl = ("%s: %d^3 MAC grid, %d FLIP particles, divergence-free curl-noise start (RMS %g m/s), red-black SOR "
              "%s, vorticity %g, currents %g, marching-tets surface with iso calibration every %d substeps, "
              "1/60 s frames; %s" % (a.config, n, nP, cfg["curl_init"], sor_txt, cfg["vorticity"], cfg["currents"],
                                     cfg["surface_every"], window))
    else:
        wl = ("%s: %d^3 MAC grid, %d FLIP particles, red-black SOR %s, vorticity %g, thermal off, "
              "back-to-back substeps; %s" % (a.config, n, nP, sor_txt, cfg["vorticity"], window))
    out = {
        "metric": "FLIP substeps/s at 128^3, 2M particles",
        "value": sps,
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * el / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded planet built by the library's world builder)",
        "config": {"workload": wl, "grid": n, "particles": nP, "particle_updates_per_s": sps * nP,
                   "pressure_ms_per_solve": g["pressureMs"] / max(g["pressureSolves"], 1),
                   "sor_sweeps_per_substep": g["sweepsTotal"] / steps,
                   "residual_checks_per_substep": g["residualsTotal"] / steps,
                   "thermal_ticks": int(g["ticksTotal"]),
                   "substep_algorithmic_GBps": b_total / el / 1e9,
                   "substep_hbm_frac": b_total / el / HBM_PEAK,
                   "render_readback_ms": readback_ms, "sor_mode": mode,
                   "autotuner": tune_report.replace("\n", "; ") if tune_report else None,
                   "learned_warm_start": learned},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / (HBM_PEAK / 1e9), "traffic": traffic,
                     "kernel": kname + kdesc,
                     "avg_launch_us": 1e6 * sor_avg_s, "algorithmic_bytes_per_launch": sor_bytes},
    }
    if not a.no_cpu_baseline:
        log("cpu baseline (oracle, 1 thread)")
        out["cpu_baseline"] = cpu_baseline(s, world, cfg)
    s.close()
    print(json.dumps(out), flush=True)


def learned_compare(s, world, cfg, weights, n, done):
    """n substeps in the library's compare mode: every pressure solve runs
    twice from the same assembled system, plain SOR from the previous
    pressure and SOR from the network's prediction, both to the tolerance.
    Returns the per-solve means. One untimed substep in compare mode comes
    first: the network's first call allocates its buffers."""
    s.load_net_weights(weights)
    s.set(pressureCompare=1)
    done += advance(s, world, cfg, 1, done)
    s.sync()
    s.set(timing=1)
    advance(s, world, cfg, n, done)
    s.sync()
    g = {k: s.get(k) for k in ("cmpMsPlain", "cmpMsLearned", "cmpSweepsPlain", "cmpSweepsLearned", "cmpSolves",
                               "netMs", "netCalls")}
    s.set(timing=0, pressureCompare=0)
    ns = max(g["cmpSolves"], 1)
    out = {"solves": int(g["cmpSolves"]), "pressure_ms_plain": g["cmpMsPlain"] / ns,
           "pressure_ms_learned": g["cmpMsLearned"] / ns, "sweeps_plain": g["cmpSweepsPlain"] / ns,
           "sweeps_learned": g["cmpSweepsLearned"] / ns, "inference_ms": g["netMs"] / max(g["netCalls"], 1),
           "speedup": g["cmpMsPlain"] / g["cmpMsLearned"] if g["cmpMsLearned"] > 0 else 0.0}
    log("learned warm start: plain %.3f ms (%.1f sweeps), learned %.3f ms (%.1f sweeps), %.2fx"
        % (out["pressure_ms_plain"], out["sweeps_plain"], out["pressure_ms_learned"], out["sweeps_learned"],
           out["speedup"]))
    return out


def learned_main(a, cfg):
    """Pressure-solve time to relative residual 1e-4 with the learned warm
    start against plain SOR, both from the same assembled system every
    substep (the library's compare mode), on the full-physics 128^3 scene."""
    if not os.path.exists(a.weights):
        print("no trained weights at %s (run watersim-v1_amd/train_warmstart.py)" % a.weights, file=sys.stderr)
        sys.exit(2)
    s, world = make_solver(cfg)
    n, nP = cfg["n"], int(s.get("nP"))
    s.load_net_weights(a.weights)
    done = advance(s, world, cfg, a.warmup, 0)
    s.sync()
    s.set(timing=1, pressureCompare=1)
    t0 = time.perf_counter()
    steps = advance(s, world, cfg, a.steps, done)
    s.sync()
    el = time.perf_counter() - t0
    g = {k: s.get(k) for k in ("cmpMsPlain", "cmpMsLearned", "cmpSweepsPlain", "cmpSweepsLearned", "cmpSolves",
                               "netMs", "netCalls")}
    s.set(timing=0, pressureCompare=0)
    ns = max(g["cmpSolves"], 1)
    plain, learned = g["cmpMsPlain"] / ns, g["cmpMsLearned"] / ns
    infer = g["netMs"] / max(g["netCalls"], 1)
    coarse = bool(s.get("netCoarse"))
    flops = net_flops(n, coarse)
    nbytes = net_bytes(n, coarse)
    tflops = flops / (infer / 1e3) / 1e12 if infer > 0 else 0.0
    # the binding roof at full speed is HBM: the bytes take longer at 8 TB/s
    # than the MFMA work at 2.5 PFLOP/s
    achieved = nbytes / (infer / 1e3) / 1e9 if infer > 0 else 0.0
    log("plain %.3f ms (%.1f sweeps), learned %.3f ms (%.1f sweeps, inference %.3f ms), %d solves"
        % (plain, g["cmpSweepsPlain"] / ns, learned, g["cmpSweepsLearned"] / ns, infer, ns))
    out = {
        "metric": "pressure-solve speed-up to residual 1e-4, learned warm start vs plain SOR, 128^3 / 2M particles",
        "value": plain / learned if learned > 0 else 0.0,
        "unit": "x",
        "n_gpus": 1, "steps": steps, "warmup": a.warmup, "ms_per_step": 1e3 * el / max(steps, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic (seeded planet built by the library's world builder; world seed held out of training)",
        "config": {"workload": "cfg4: the cfg3 full-physics scene, %d^3, %d particles; every pressure solve run "
                               "twice from the same system: plain red-black SOR from the previous pressure, and "
                               "SOR from the network's prediction, both to relative residual 1e-4; the "
                               "network runs on the %s grid" % (n, nP, "coarsened (%d^3)" % (n // 2) if coarse
                                                                else "full"),
                   "net_coarse": int(coarse),
                   "pressure_ms_plain": plain, "pressure_ms_learned": learned,
                   "sweeps_plain": g["cmpSweepsPlain"] / ns, "sweeps_learned": g["cmpSweepsLearned"] / ns,
                   "inference_ms": infer, "solves": int(ns)},
        "roofline": {"bound": "hbm" if nbytes / HBM_PEAK >= flops / BF16_PEAK else "mfma",
                     "achieved": achieved, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / (HBM_PEAK / 1e9), "traffic": None,
                     "kernel": "warm-start network inference (BF16 MFMA implicit-GEMM convolutions, LDS-tiled or "
                               "split-K by level size, + elementwise)",
                     "avg_launch_us": 1e3 * infer, "algorithmic_bytes_per_launch": nbytes,
                     "algorithmic_flops_per_launch": flops, "mfma_tflops": tflops,
                     "mfma_frac": tflops / (BF16_PEAK / 1e12)},
    }
    s.close()
    print(json.dumps(out), flush=True)


def make_planets(cfg, weights, only=None):
    """The batched-planets workload: cfg["planets"] planets whose grid sides
    are drawn from cfg["sizes"] (seeded), each with its own terrain seed and
    gravity (0.8..1.2 g), about 8 particles per liquid cell, the learned warm
    start and SOR to the tolerance. only: indices of the planets to build
    (all by default). Returns the solvers and every planet's grid side."""
    from concurrent.futures import ThreadPoolExecutor
    import watersim_hip
    rng = np.random.default_rng(cfg["seed"])
    P = cfg["planets"]
    sizes = [int(x) for x in rng.choice(cfg["sizes"], size=P)]
    gravs = [float(9.8 * rng.uniform(0.8, 1.2)) for _ in range(P)]
    idx = list(range(P)) if only is None else list(only)
    t = time.time()

    def build(i):
        n = sizes[i]
        return watersim_hip.build_world(n, W_DOMAIN / n, 0.25 * W_DOMAIN, 0.15 * W_DOMAIN, cfg["seed"] + 1 + i,
                                        int(cfg["per_cell"] * n ** 3))
    with ThreadPoolExecutor(8) as ex:
        worlds = list(ex.map(build, idx))
    log("%d worlds built in %.1fs" % (len(idx), time.time() - t))
    planets = []
    for i, w in zip(idx, worlds):
        s = watersim_hip.solver_for_world(w)
        s.set(pic=0.05, omega=1.7, gravity=gravs[i], maxSpeed=3.2, vorticity=cfg["vorticity"], cfl=3.0,
              pressureTol=cfg["tol"], maxPressureIters=4000, residualEvery=8)
        s.load_net_weights(weights)
        s.set(warmStart=1)
        planets.append(s)
    return planets, sizes


def batched_main(a, cfg):
    """Batched planets: independent planets of mixed sizes advanced together
    by the library's batch (ws_batch_*) on one stream: each stage is one
    launch over all planets (the planets' own stages merged from their
    recorded launches, the warm start's network steps, all pressure solves
    to 1e-4 with the residual checks on the device, the extrapolation by
    tile diagonal), and the whole batched substep is captured into a HIP
    graph and replayed. The autotuner picks the P2G bin per grid size (on
    one planet of each size), then the convolution form, the P2G bin and
    the solve's workgroup size for the batch as a whole.
    value = particle updates per second summed over the planets."""
    import watersim_hip
    if not os.path.exists(a.weights):
        raise SystemExit("cfg5 needs the trained warm-start weights (%s)" % a.weights)
    planets, sizes = make_planets(cfg, a.weights)
    P = len(planets)
    nPs = [int(s.get("nP")) for s in planets]
    log("%d planets, %d particles (%d .. %d)" % (P, sum(nPs), min(nPs), max(nPs)))
    batch = watersim_hip.Batch(planets, graph=True, threads=1024)
    batch.run(a.warmup)
    # autotuner: convolution tiling and P2G bin per grid size (one planet of
    # each size), then the solve's workgroup size for the whole batch
    tiling, bins = {}, {}
    for n in sorted(set(sizes)):
        s = planets[sizes.index(n)]
        rep = s.tune(64)
        tiling[n] = int(s.get("netTiled"))
        bins[n] = int(s.get("p2gBin"))
        log("tune %d^3: %s" % (n, rep.replace("\n", "; ")))
    for i, s in enumerate(planets):
        s.set(netTiled=tiling[sizes[i]], p2gBin=bins[sizes[i]])
    # the convolutions' form for the batch as a whole: each size's own choice
    # (a single planet's grid may be too small to fill the GPU with LDS
    # tiles), or LDS tiles on every grid (the planets fill it together)
    net_trial = {}
    for form in ("per size", 2):
        for i, s in enumerate(planets):
            s.set(netTiled=tiling[sizes[i]] if form == "per size" else form)
        t1 = time.perf_counter()
        batch.run(10)
        net_trial[form] = (time.perf_counter() - t1) / 10
        log("batch with netTiled %s: %.3f ms per batched substep" % (form, 1e3 * net_trial[form]))
    net_form = min(net_trial, key=net_trial.get)
    for i, s in enumerate(planets):
        s.set(netTiled=tiling[sizes[i]] if net_form == "per size" else net_form)
    # the P2G bin for the batch as a whole: each size's own choice (small
    # bins keep one planet's GPU busy), or one bin for every planet (larger
    # bins visit each particle from fewer tiles; the planets fill the GPU
    # together)
    bin_trial = {}
    for choice in ["per size"] + list(range(4)):
        for i, s in enumerate(planets):
            s.set(p2gBin=bins[sizes[i]] if choice == "per size" else choice)
        t1 = time.perf_counter()
        batch.run(10)
        bin_trial[choice] = (time.perf_counter() - t1) / 10
        log("batch with p2gBin %s: %.3f ms per batched substep" % (choice, 1e3 * bin_trial[choice]))
    bin_choice = min(bin_trial, key=bin_trial.get)
    for i, s in enumerate(planets):
        s.set(p2gBin=bins[sizes[i]] if bin_choice == "per size" else bin_choice)
    # the solve's workgroup size, and each planet's workgroups on one XCD
    # (L2-local barriers) or dealt over all XCDs (agent-scope barriers)
    trial = {}
    for xl in (1, 0):
        for th in (256, 512, 1024):
            batch.set(threads=th, xcdLocal=xl)
            t1 = time.perf_counter()
            batch.run(10)
            trial[(th, xl)] = (time.perf_counter() - t1) / 10
            log("batch with %d-thread solves, xcdLocal %d (%d planets XCD-local): %.3f ms per batched substep" % (
                th, xl, int(batch.get("xcdLocalPlanets")), 1e3 * trial[(th, xl)]))
    threads, xcd_local = min(trial, key=trial.get)
    batch.set(threads=threads, xcdLocal=xcd_local)
    log("timing %d batched substeps (HIP graph), %d-thread solves, xcdLocal %d" % (a.steps, threads, xcd_local))
    t0 = time.perf_counter()
    done = 0
    while done < a.steps:
        k = min(100, a.steps - done)
        batch.run(k)
        done += k
        log("  %d substeps per planet" % done)
    el = time.perf_counter() - t0
    updates = sum(nPs) * a.steps
    log("aggregate %.3g particle-updates/s, wall %.1f s" % (updates / el, el))
    # the dominant kernel, the joint solve, timed with HIP events on the batch
    # stream over a few eager substeps
    batch.set(timing=1)
    batch.run(4)
    g = {k: batch.get(k) for k in ("solveMs", "solveLaunches", "cellSweeps", "sweeps")}
    local_planets = int(batch.get("xcdLocalPlanets"))
    batch.set(timing=0)
    launches = max(g["solveLaunches"], 1)
    solve_s = g["solveMs"] / 1e3 / launches
    solve_b = 13.0 * g["cellSweeps"] / launches
    achieved = solve_b / solve_s / 1e9 if g["solveMs"] > 0 else 0.0
    sweeps_per = g["sweeps"] / launches / P
    per_size = {}
    for n in sorted(set(sizes)):
        idx = [i for i in range(P) if sizes[i] == n]
        per_size["%d^3" % n] = {"planets": len(idx), "particles_mean": float(np.mean([nPs[i] for i in idx])),
                                "netTiled": tiling[n] if net_form == "per size" else net_form,
                                "p2gBin": bins[n] if bin_choice == "per size" else bin_choice}
    out = {
        "metric": "aggregate particle-updates/s, %d batched planets (32^3..96^3, learned warm start + SOR to 1e-4)" % P,
        "value": updates / el,
        "unit": "particle-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * el / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (planets built by the library's world builder, seeds %d..%d)" % (cfg["seed"] + 1,
                                                                                            cfg["seed"] + P),
        "config": {"workload": "cfg5: %d independent planets, grid sizes drawn from %s, about 8 particles per "
                               "liquid cell, own terrain seed and gravity each, learned warm start + red-black SOR "
                               "to 1e-4, vorticity %g; one batch on one stream: every stage one launch over all "
                               "planets (the planets' own stages merged from their recorded launches, the "
                               "network's steps, all pressure solves with workgroups per planet by size, the "
                               "extrapolation by tile diagonal), the batched substep replayed from a HIP graph; "
                               "convolution form, P2G bin and solve workgroup size autotuned for the batch"
                               % (P, list(cfg["sizes"]), cfg["vorticity"]),
                   "particles_total": sum(nPs), "per_size": per_size,
                   "planet_substep_ms": 1e3 * el / a.steps,
                   "solve_threads": threads, "xcd_local": xcd_local, "xcd_local_planets": local_planets,
                   "batch_trial_ms": {"%d threads, xcdLocal %d" % k: 1e3 * v for k, v in trial.items()},
                   "net_form_trial_ms": {"netTiled %s" % k: 1e3 * v for k, v in net_trial.items()},
                   "p2g_bin_trial_ms": {"p2gBin %s" % k: 1e3 * v for k, v in bin_trial.items()},
                   "stage_launches": int(batch.get("stageLaunches")),
                   "joint_solve_ms": 1e3 * solve_s, "sweeps_per_planet_substep": sweeps_per},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / (HBM_PEAK / 1e9), "traffic": None,
                     "kernel": "k_sor_solve_batch (every planet's whole pressure solve, workgroups per planet by size)",
                     "avg_launch_us": 1e6 * solve_s, "algorithmic_bytes_per_launch": solve_b},
    }
    batch.close()
    for s in planets:
        s.close()
    print(json.dumps(out), flush=True)


def read_traffic(kernel, config):
    """HBM bytes per launch of the kernel from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, copy-kernel calibrated): the
    line 'traffic_bytes_per_launch: <bytes>' of profiles/<kernel>_traffic_<config>.md."""
    path = os.path.join(ROOT, "profiles", "%s_traffic_%s.md" % (kernel, config))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.strip().startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


if __name__ == "__main__":
    main()
