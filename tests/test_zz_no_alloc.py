"""The allocation-free dense eigensolver (bse_eigh_dense_noalloc) on the MI355X.

It runs the ordinary pipeline with every buffer cut from a fixed device
array, so its results must equal those of bse_eigh_dense on the same matrix:
eigenvalues to 1e-12 ||A||, eigenvectors as subspaces per cluster (residual
and orthonormality at FP64 level).  Two calls in a row, the second with a
smaller matrix (the array keeps the first call's contents), plus a
two-stage-sized call.  Inputs and outputs are device buffers allocated by the
test through the HIP runtime the library itself loaded (ctypes; no torch,
whose bundled runtime must not share the process with another one).
"""
import ctypes as C

import numpy as np
import pytest

from bse_amd import _abi
from parity_util import cluster_overlap_error

pytestmark = pytest.mark.gpu


def _lib():
    lib = _abi.load_library()
    lib.bse_eigh_dense_noalloc.restype = C.c_int
    lib.bse_eigh_dense_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_int, C.c_int,
                                           C.c_void_p, C.c_void_p, C.c_int64]
    lib.bse_noalloc_info.restype = C.c_int
    lib.bse_noalloc_info.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    return lib


def _herm(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return np.asfortranarray(0.5 * (a + a.conj().T))


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime the library loaded."""

    def __init__(self):
        _abi.load_library()
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 256))) == 0
        return p.value

    def h2d(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        assert self.rt.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1) == 0

    def d2h(self, arr, src):
        assert self.rt.hipMemcpy(arr.ctypes.data, src, arr.nbytes, 2) == 0
        return arr

    def free(self, p):
        assert self.rt.hipFree(p) == 0

    def sync(self):
        assert self.rt.hipDeviceSynchronize() == 0


def _noalloc(ctx, lib, hip, a, k):
    n = a.shape[0]
    nz = n if k < 0 else k
    ad = hip.malloc(16 * n * n)
    ev = hip.malloc(8 * n)
    zd = hip.malloc(16 * n * max(nz, 1))
    try:
        hip.h2d(ad, np.asfortranarray(a).ravel(order="F"))
        rc = lib.bse_eigh_dense_noalloc(ctx.h, n, ad, n, 0, k, ev, zd, n)
        if rc != 0:
            raise _abi.BSEError(f"noalloc failed ({rc}): {lib.bse_last_error(ctx.h).decode()}")
        hip.sync()
        e = hip.d2h(np.zeros(n), ev)
        z = hip.d2h(np.zeros(n * max(nz, 1), dtype=np.complex128), zd).reshape(max(nz, 1), n).T[:, :nz]
    finally:
        for p in (ad, ev, zd):
            hip.free(p)
    return e, z


def test_noalloc_matches_dense_entry_twice_then_smaller():
    lib = _lib()
    hip = _Hip()
    with _abi.Context(0) as ctx:
        info = (C.c_int64 * 3)()
        assert lib.bse_noalloc_info(ctx.h, info) == 0
        assert info[1] == 512
        for n, k, seed in [(512, -1, 1), (200, 16, 2), (300, 8, 3)]:
            print(f"noalloc n={n} k={k}", flush=True)
            a = _herm(n, seed)
            ev0, z0 = ctx.eigh_dense(a, "fp64", k)
            ev1, z1 = _noalloc(ctx, lib, hip, a, k)
            scale = np.abs(ev0).max()
            assert np.abs(ev1 - ev0).max() < 1e-12 * scale
            kk = z0.shape[1]
            assert cluster_overlap_error(z1, z0, ev0, 1e-10 * scale, ncol=kk) < 1e-10
            assert np.abs(a @ z1 - z1 * ev1[None, :kk]).max() < 1e-11 * scale * np.sqrt(n)
            assert np.abs(z1.conj().T @ z1 - np.eye(kk)).max() < 1e-10
            assert lib.bse_noalloc_info(ctx.h, info) == 0
            print(f"  scratch used {info[2] / 2**20:.1f} MiB of {info[0] / 2**20:.0f}", flush=True)
            assert 0 < info[2] <= info[0]


def test_noalloc_rejects_too_large():
    lib = _lib()
    hip = _Hip()
    with _abi.Context(0) as ctx:
        a = hip.malloc(16 * 600 * 600)
        ev = hip.malloc(8 * 600)
        try:
            rc = lib.bse_eigh_dense_noalloc(ctx.h, 600, a, 600, 0, 0, ev, None, 600)
        finally:
            hip.free(a)
            hip.free(ev)
        assert rc == -1
