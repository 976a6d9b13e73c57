"""The allocation-free product mpgpu_avprod_noalloc (no counterpart in the reference; for callers
that cannot allocate device memory) against mpgpu_avprod on the same matrix, through the C-ABI:
x and y are device arrays the caller owns.  Both entry points run the same deterministic kernels,
so the products agree bit for bit.  Two calls in a row share the fixed device scratch: the
second, on a smaller matrix, must not see the first one's partial sums.  A null pointer and
sparse storage are refused with exit code 30 before any kernel runs."""
import ctypes as C

import numpy as np
import pytest

import mpgpu

pytestmark = pytest.mark.gpu


def _full_problem(seed, nl, nx, ny, ntr):
    rp, ii, gg, _ = mpgpu.dense_records(seed, nl, nx, ny, ntr)
    P = mpgpu.Pede(matsto=mpgpu.FULL, metsol=mpgpu.DECOMPOSITION, mreqenf=1, mcount=1)
    P.add_records(rp, ii, gg)
    P.build_structure()
    P.loopn()
    return P


class _Dev:
    """Device arrays of the caller, from the HIP runtime the library itself uses."""

    def __init__(self):
        import os
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        self.hip = C.CDLL(path)
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipDeviceSynchronize.argtypes = []

    def alloc(self, n):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), 8 * n) == 0
        return p

    def put(self, p, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), 8 * len(a), 1) == 0   # host -> device

    def get(self, p, n):
        a = np.empty(n)
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, 8 * n, 2) == 0       # device -> host
        return a

    def free(self, p):
        self.hip.hipFree(p)


def _noalloc(P, x):
    d = _Dev()
    xd, yd = d.alloc(len(x)), d.alloc(len(x))
    d.put(xd, x)
    d.put(yd, np.full(len(x), np.nan))
    rc = P.L.mpgpu_avprod_noalloc(P.h, xd, yd)
    y = d.get(yd, len(x))
    d.free(xd)
    d.free(yd)
    return rc, y


def test_avprod_noalloc_matches_avprod():
    rng = np.random.default_rng(3)
    sizes = []
    for (nl, nx, ny) in [(12, 8, 5), (5, 3, 2)]:        # 2880 then 180 rows
        P = _full_problem(41, nl, nx, ny, 4000)
        x = rng.normal(size=P.nall)
        y_ref = P.avprod(x)
        rc, y = _noalloc(P, x)
        assert rc == 0, P.L.mpgpu_last_error(P.h)
        assert np.array_equal(y, y_ref)
        sizes.append(P.nall)
        P.close()
    assert sizes[1] < sizes[0]


def test_avprod_noalloc_refusals():
    P = _full_problem(42, 3, 2, 2, 500)
    d = _Dev()
    xd = d.alloc(P.nall)
    assert P.L.mpgpu_avprod_noalloc(P.h, None, xd) == 30
    rp, ii, gg, _ = mpgpu.dense_records(43, 3, 2, 2, 500)
    S = mpgpu.Pede(matsto=mpgpu.SPARSE, metsol=mpgpu.MINRES, mreqenf=1, mcount=1)
    S.add_records(rp, ii, gg)
    S.build_structure()
    S.loopn()
    rc = S.L.mpgpu_avprod_noalloc(S.h, xd, xd)
    assert rc == 30 and b"full matrix storage" in S.L.mpgpu_last_error(S.h)
    d.free(xd)
