"""lddmm_multiply_kernel_cross_noalloc: the K.alpha operator on device arrays
with its partial sums in the library's fixed per-GPU scratch, called through
the C-ABI, against lddmm_multiply_kernel_cross (which allocates its own
device buffers) on the same inputs: bitwise equal, on two calls in a row,
the second smaller than the first (the scratch keeps the first call's
partials, which the second must not read)."""
import ctypes as C

import numpy as np
import pytest

from lddmm_amd import geometry as G

pytestmark = pytest.mark.gpu


def _dev(L, arr):
    arr = np.ascontiguousarray(arr, dtype=np.float64)
    p = C.c_void_p()
    assert L.lddmm_device_malloc(C.byref(p), max(arr.nbytes, 8)) == 0
    if arr.nbytes:
        assert L.lddmm_memcpy_htod(p, arr.ctypes.data, arr.nbytes) == 0
    return p


def _noalloc(gpu, qi, qj, a, order, width):
    """Upload (the test allocates; the entry point does not), call, download."""
    L = gpu.lib()
    ni, d = qi.shape
    nj = qj.shape[0]
    bufs = [_dev(L, qi.T), _dev(L, qj.T), _dev(L, a.T)]
    out = C.c_void_p()
    assert L.lddmm_device_malloc(C.byref(out), max(ni * d * 8, 8)) == 0
    L.lddmm_multiply_kernel_cross_noalloc.argtypes = [C.c_void_p] * 4 + [C.c_int, C.c_double, C.c_int, C.c_int,
                                                                          C.c_int, C.c_void_p]
    rc = L.lddmm_multiply_kernel_cross_noalloc(out, bufs[0], bufs[1], bufs[2], order, width, ni, nj, d, None)
    err = L.lddmm_last_error().decode() if rc else ""
    assert L.lddmm_device_synchronize() == 0
    res = np.empty((d, ni))
    assert L.lddmm_memcpy_dtoh(res.ctypes.data, out, ni * d * 8) == 0
    for b in bufs + [out]:
        L.lddmm_device_free(b)
    return rc, err, res.T.copy()


@pytest.mark.parametrize("dim,order", [(3, -1), (3, 1), (2, -1), (2, 2)])
def test_noalloc_multiply_kernel_equals_allocating_entry(gpu, dim, order):
    for call, (ni, nj) in enumerate(((3000, 5000), (700, 257))):
        qi = (0.5 * G.normal(10 + call, ni * dim)).reshape(ni, dim)
        qj = (0.5 * G.normal(20 + call, nj * dim)).reshape(nj, dim)
        a = G.normal(30 + call, nj * dim).reshape(nj, dim)
        print("noalloc K.a dim=%d order=%d call %d: %d x %d" % (dim, order, call, ni, nj), flush=True)
        rc, err, got = _noalloc(gpu, qi, qj, a, order, 0.3)
        assert rc == 0, err
        ref = gpu.multiplyKernel(qi, qj, a, order, 0.3)
        assert np.array_equal(got, ref)


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    """2^21 doubles of partials: 800,000 targets in 3D need 2.4e6 -> -1."""
    L = gpu.lib()
    L.lddmm_multiply_kernel_cross_noalloc.argtypes = [C.c_void_p] * 4 + [C.c_int, C.c_double, C.c_int, C.c_int,
                                                                          C.c_int, C.c_void_p]
    p = C.c_void_p()
    assert L.lddmm_device_malloc(C.byref(p), 64) == 0
    # the size check comes before any launch: the small buffer is never read
    rc = L.lddmm_multiply_kernel_cross_noalloc(p, p, p, p, -1, 0.3, 800000, 1000, 3, None)
    msg = L.lddmm_last_error()
    L.lddmm_device_free(p)
    assert rc == -1 and b"scratch" in msg
