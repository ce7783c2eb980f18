"""ws_mt_build_noalloc (marching tetrahedra with no allocation: caller device
buffers, the library's static scratch) against ws_mt_build on the same
fields: vertex count, FP64 volume, positions and normals bit for bit, over
calls in a row on fields of decreasing size (the scratch keeps the previous
call's contents), with a vertex cap that cuts a triangle, on the context's
own density field, and the refusal of fields above the scratch's size."""
import ctypes

import numpy as np
import pytest

from gpu_scenes import watersim_hip

pytestmark = pytest.mark.gpu

HIP = None


def hip():
    global HIP
    if HIP is None:
        HIP = ctypes.CDLL("libamdhip64.so")
        HIP.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        HIP.hipFree.argtypes = [ctypes.c_void_p]
        HIP.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        HIP.hipDeviceSynchronize.argtypes = []
    return HIP


class DevBuf:
    """A caller-owned device buffer (allocated by the test, not the library)."""

    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.ptr = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.ptr), max(nbytes, 4)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert hip().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0  # host to device

    def get(self, n, dtype):
        a = np.empty(n, dtype)
        assert hip().hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.ptr, a.nbytes, 2) == 0  # device to host
        return a

    def free(self):
        hip().hipFree(self.ptr)


def blobs(n, seed, k=6):
    """A smooth corner field on (n+1)^3 points: a sum of Gaussian bumps."""
    rng = np.random.default_rng(seed)
    g = np.arange(n + 1, dtype=np.float64) / n
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    f = np.zeros_like(x)
    for _ in range(k):
        c = rng.uniform(0.2, 0.8, 3)
        r = rng.uniform(0.08, 0.2)
        f += 2.5 * np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (r * r))
    return f.astype(np.float32).ravel()


def compare(s, field, n, dx, iso, cap=None):
    cnt, vol, pos, nrm = s.mt_build(dens=field, dims=(n, n, n), dx=dx, iso=iso)
    assert cnt > 0
    cap = cnt if cap is None else cap
    d = DevBuf(field.nbytes)
    d.put(field)
    p, q = DevBuf(12 * cap), DevBuf(12 * cap)
    cnt2, vol2 = s.mt_build_noalloc(d.ptr.value, (n, n, n), dx, iso, p.ptr.value, q.ptr.value, cap)
    got_p, got_n = p.get(3 * cap, np.float32), q.get(3 * cap, np.float32)
    print("  %d^3 field: %d vertices (cap %d), volume %.12g / %.12g" % (n, cnt, cap, vol, vol2))
    assert cnt2 == cnt
    assert vol2 == vol
    assert np.array_equal(got_p, pos[:3 * cap])
    assert np.array_equal(got_n, nrm[:3 * cap])
    # count and volume only
    cnt3, vol3 = s.mt_build_noalloc(d.ptr.value, (n, n, n), dx, iso)
    assert (cnt3, vol3) == (cnt, vol)
    for b in (d, p, q):
        b.free()


def test_noalloc_equals_mt_build_on_consecutive_calls():
    s = watersim_hip.Solver(48, 48, 48, 0.1, 16)
    compare(s, blobs(48, 1), 48, 0.1, 1.0)
    # a smaller field next, with a cap that cuts a triangle
    small = blobs(20, 2, k=3)
    cnt = s.mt_build(dens=small, dims=(20, 20, 20), dx=0.25, iso=0.9)[0]
    compare(s, small, 20, 0.25, 0.9, cap=cnt - 4)
    # the context's own density field
    own = blobs(48, 3)
    s.upload("dens", own)
    cnt, vol, pos, nrm = s.mt_build(dims=(48, 48, 48), dx=0.1, iso=1.2)
    p, q = DevBuf(12 * cnt), DevBuf(12 * cnt)
    cnt2, vol2 = s.mt_build_noalloc(None, (48, 48, 48), 0.1, 1.2, p.ptr.value, q.ptr.value, cnt)
    assert (cnt2, vol2) == (cnt, vol)
    assert np.array_equal(p.get(3 * cnt, np.float32), pos)
    assert np.array_equal(q.get(3 * cnt, np.float32), nrm)
    p.free()
    q.free()
    s.close()


def test_noalloc_refuses_fields_above_its_scratch():
    s = watersim_hip.Solver(8, 8, 8, 0.1, 16)
    n = 129  # 129^3 cells > 2^21
    with pytest.raises(watersim_hip.WatersimError) as e:
        s.mt_build_noalloc(None, (n, n, n), 0.1, 1.0)
    assert "at most" in str(e.value)
    s.close()
