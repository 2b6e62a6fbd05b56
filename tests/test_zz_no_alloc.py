"""The no-allocation entry point andr_bicgstab_dense_noalloc against andr_bicgstab_dense.

The test allocates the device arrays itself through the HIP runtime the library is linked with
(the entry point must not allocate). Two calls in a row on the same device, the second with a
smaller system and fewer right-hand sides, so the shared fixed work space is reused with another
layout; each result must equal the allocating entry point's on the same inputs (same kernels and
summation order: the test asks for agreement to 1e-13 relative and reports whether the two are
bit-identical). A system whose work space does not fit the fixed array is refused with
ANDR_ERR_ARG before anything runs.
"""
import ctypes as C

import numpy as np
import pytest

from andromeda_amd import api

pytestmark = pytest.mark.gpu


class _Dev:
    """Device buffers from the HIP runtime of libandromeda_hip (test infrastructure)."""

    def __init__(self):
        api.load_library()
        self.hip = C.CDLL("libamdhip64.so.7")  # the runtime instance the library already loaded
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def get(self, ptr, shape):
        out = np.empty(shape)
        assert self.hip.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        self.hip.hipDeviceSynchronize()
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _system(n, nrhs, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) / np.sqrt(n) + 3.0 * np.eye(n)
    B = rng.normal(size=(n, nrhs))
    X0 = 0.01 * rng.normal(size=(n, nrhs))
    return A, B, X0


def test_noalloc_matches_allocating_entry_point_twice(ctx):
    dev = _Dev()
    try:
        for n, nrhs, pad, seed in ((3000, 11, 16, 1), (777, 3, 5, 2)):
            A, B, X0 = _system(n, nrhs, seed)
            Xr, ir = ctx.bicgstab_dense(A, B, X0=X0, tol=1e-12, maxit=500)
            lda = n + pad
            Ap = np.zeros((n, lda))
            Ap[:, :n] = A
            dA, dB, dX = dev.put(Ap), dev.put(B.T), dev.put(X0.T)  # B, X: column k contiguous
            inn = ctx.bicgstab_dense_noalloc(n, dA, lda, nrhs, dB, dX, tol=1e-12, maxit=500)
            Xn = dev.get(dX, (nrhs, n)).T
            rel = np.abs(Xn - Xr).max() / np.abs(Xr).max()
            print(f"n={n} nrhs={nrhs}: rel diff {rel:.2e}, bit-identical {np.array_equal(Xn, Xr)}, its {inn['nit']}")
            assert inn["nit"] == ir["nit"] and inn["status"] == ir["status"] == [0] * nrhs
            assert rel < 1e-13, rel
            res = np.linalg.norm(A @ Xn - B) / np.linalg.norm(B)
            assert res < 1e-11, res
    finally:
        dev.free()


def test_noalloc_refuses_oversize(ctx):
    dev = _Dev()
    try:
        d = dev.put(np.zeros(16))
        with pytest.raises(api.AndromedaError, match=r"\(-1\).*work space"):
            ctx.bicgstab_dense_noalloc(40000, d, 40000, 1, d, d)
    finally:
        dev.free()
