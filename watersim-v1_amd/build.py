"""Builds the watersim HIP library (gfx950) in-tree,
watersim-v1_amd/_build/libwatersim_hip.so, and the Node N-API addon over its
C ABI, watersim-v1_amd/_build/watersim_napi.node (when Node's headers are
present)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = ["ws_grid.hip", "ws_particles.hip", "ws_surface.hip", "ws_api.hip", "ws_api_surface.hip", "ws_thermal.hip", "ws_frame.hip", "ws_learn.hip", "ws_contacts.hip", "ws_balls.hip", "ws_metaballs.hip", "ws_batch.hip", "ws_noalloc.hip", "ws_sor_xcd.hip", "ws_extrap.hip", "ws_scan.hip", "ws_world.cpp"]
OUT = os.path.join(HERE, "_build", "libwatersim_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# per-file additions: the ball, contact, metaball and surface code evaluate the reference's double
# formulas operation by operation, so nothing may fuse into an FMA
FILE_FLAGS = {}
# -ffp-contract=off: no multiply-add is fused unless the source asks for it,
# so the bits of a computation follow from its source alone, whichever
# kernel it is inlined into (a stage's body runs in its own kernel and in
# the kernel that runs it for many planets at once)
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics", "-ffp-contract=off"]


def build(verbose=False):
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    objs = []
    procs = []
    for s in SRC:
        src = os.path.join(HERE, "csrc", s)
        obj = os.path.join(HERE, "_build", os.path.splitext(s)[0] + ".o")
        objs.append(obj)
        cmd = [HIPCC] + FLAGS + FILE_FLAGS.get(s, []) + ["-c", src, "-o", obj]
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
    build_addon()
    return OUT


NODE_INC = "/usr/include/node"
ADDON = os.path.join(HERE, "_build", "watersim_napi.node")


def build_addon():
    """The N-API addon: plain C against node_api.h, linked to the library
    with an $ORIGIN rpath so the two travel together."""
    if not os.path.exists(os.path.join(NODE_INC, "node_api.h")):
        return None
    cmd = ["gcc", "-O2", "-std=c11", "-shared", "-fPIC", "-Wall", "-Werror", "-DNAPI_VERSION=8",
           "-DNODE_GYP_MODULE_NAME=watersim_napi", "-I" + NODE_INC, "-I" + os.path.join(HERE, "..", "include"),
           os.path.join(HERE, "addon", "watersim_napi.c"), "-o", ADDON,
           "-L" + os.path.join(HERE, "_build"), "-lwatersim_hip", "-Wl,-rpath,$ORIGIN"]
    subprocess.run(cmd, check=True)
    return ADDON


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
