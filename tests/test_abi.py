"""CPU-side checks of the C ABI: the HIP library loads and exports exactly the
entry points include/watersim.h declares (no device calls)."""
import ctypes
import os
import re
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "watersim-v1_amd"))

import watersim_hip  # noqa: E402


def declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for h in sorted(os.listdir(inc)):
        if h.endswith(".h"):
            txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(inc, h)).read(), flags=re.S)
            names |= set(re.findall(r"\b(ws_[a-z_0-9]+)\s*\(", txt))
    return sorted(names)


def test_header_matches_binding_list():
    assert declared() == sorted(watersim_hip.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = watersim_hip.load()
    for name in declared():
        assert hasattr(L, name), name
        assert ctypes.cast(getattr(L, name), ctypes.c_void_p).value


def test_oracle_is_not_linked_into_the_library():
    # the product path must not reach the oracle
    import subprocess
    out = subprocess.run(["nm", "-D", watersim_hip.LIB_PATH], capture_output=True, text=True).stdout
    syms = [l.split()[-1] for l in out.splitlines() if l.strip()]
    assert not [x for x in syms if x.startswith("oc_")]
