"""Builds the watersim HIP library (gfx950) in-tree:
watersim-v1_amd/_build/libwatersim_hip.so."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = ["ws_grid.hip", "ws_particles.hip", "ws_surface.hip", "ws_api.hip", "ws_api_surface.hip", "ws_thermal.hip", "ws_frame.hip", "ws_world.cpp"]
OUT = os.path.join(HERE, "_build", "libwatersim_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def build(verbose=False):
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    objs = []
    procs = []
    for s in SRC:
        src = os.path.join(HERE, "csrc", s)
        obj = os.path.join(HERE, "_build", os.path.splitext(s)[0] + ".o")
        objs.append(obj)
        cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0 or verbose:
            sys.stderr.write(out.decode(errors="replace"))
        failed |= p.returncode != 0
    if failed:
        raise RuntimeError("hipcc failed")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", OUT] + objs
    subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
