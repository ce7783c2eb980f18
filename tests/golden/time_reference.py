"""Times the reference's CPU substep in Node (time_reference.js) and the
oracle, this repository's single-threaded C restatement of the same path
("the twin"), on identical workloads in this container, and writes
tests/golden/reference_timing.json. The ratio twin/reference lets a
measurement of the twin on another machine (the benchmark's CPU baseline)
stand for the reference's own time there.

    python tests/golden/time_reference.py /root/reference

Both sides start from the same world (the library's builder reproduces the
reference's terrain and fill bit for bit), the same velocities and the same
settings, and both time 90 substeps after 30 untimed ones, dt from the CFL
law. cfg2's start velocities are the benchmark's curl-noise field.
Run only in the development container (the reference is read from its
checkout); the GPU is not used.
"""
import ctypes
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "watersim-v1_amd")]

import bench  # noqa: E402
import cpu_oracle  # noqa: E402
import watersim_hip  # noqa: E402

W = bench.W_DOMAIN
SPECS = {"cfg1": dict(n=32, target=10000, seed=20260114, vort=0.0, tol=0.0),
         "cfg2": dict(n=64, target=232000, seed=20260115, vort=0.05, tol=1e-4)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def twin(name, world, vel):
    sp = SPECS[name]
    n = sp["n"]
    st = cpu_oracle.OracleState({"nx": n, "ny": n, "nz": n, "dx": W / n, "coreR": world["coreR"],
                                 "oceanR": world["oceanR"], "spacing": world["spacing"],
                                 "particleVolume": world["particleVolume"], "dv": world["dv"],
                                 "tn": world["tn"]})
    st.set_static(world["solid"], world["R"], world["staticTypes"], world["gu"], world["gv"], world["gw"])
    nP = world["nP"]
    parts = {"px": world["px"], "py": world["py"], "pz": world["pz"],
             "pT": np.full(nP, 0.32, np.float32)}
    if vel is not None:
        parts.update(pvx=vel[0], pvy=vel[1], pvz=vel[2])
    st.set_particles(nP, **parts)
    st.set_params(pic=0.05, omega=1.7, gravity=9.8, maxSpeed=3.2, vorticity=sp["vort"], cfl=3.0, gRamp=0.0,
                  umax=0.0)
    if sp["tol"] > 0:
        st.set_params(pressureTol=sp["tol"], residualEvery=8, maxPressureIters=4000, sorRedBlack=0)
    else:
        st.set_params(pressureIters=40, sorRedBlack=0)
    for _ in range(30):
        st.call("oc_substep", ctypes.c_double(st.call("oc_dt_law")))
    times, sweeps = [], 0
    for _ in range(90):
        dt = st.call("oc_dt_law")
        t0 = time.perf_counter()
        st.call("oc_substep", ctypes.c_double(dt))
        times.append((time.perf_counter() - t0) * 1e3)
        sweeps += st.get_i("itersDone") if sp["tol"] > 0 else 40
    times.sort()
    return {"median_ms": times[len(times) // 2], "mean_ms": sum(times) / len(times), "sweeps_per_substep": sweeps / 90,
            "ke_end": st.call("oc_kinetic_energy")}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = {"cpu": cpu_model(), "threads": 1, "configs": {}}
    for name, sp in SPECS.items():
        n = sp["n"]
        world = watersim_hip.build_world(n, W / n, 0.25 * W, 0.15 * W, sp["seed"], sp["target"])
        args = ["node", os.path.join(HERE, "time_reference.js"), ref, name]
        vel = None
        tmp = None
        if name == "cfg2":
            vel = bench.curl_noise_velocity(world["px"].astype(np.float64), world["py"].astype(np.float64),
                                            world["pz"].astype(np.float64), sp["seed"], 0.5, W)
            tmp = tempfile.NamedTemporaryFile(suffix=".bin", delete=False)
            tmp.write(np.ascontiguousarray(vel, np.float32).tobytes())
            tmp.close()
            args.append(tmp.name)
        r = json.loads(subprocess.run(args, check=True, capture_output=True, text=True).stdout.strip().splitlines()[-1])
        if tmp:
            os.unlink(tmp.name)
        assert r["particles"] == world["nP"], (r["particles"], world["nP"])
        t = twin(name, world, vel)
        out["node"] = r["node"]
        out["configs"][name] = {"particles": world["nP"], "reference": r, "twin": t,
                                "twin_over_reference": t["median_ms"] / r["median_ms"]}
        print(name, json.dumps(out["configs"][name]), flush=True)
    with open(os.path.join(HERE, "reference_timing.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
