"""The allocation-free FP64 solve (pssfss_lu_solve_batch_noalloc) against the
allocating entry point (pssfss_lu_solve_batch), through the C ABI, on two calls
in a row (the second smaller, reusing the same device scratch array)."""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _say(msg):
    print(f"\n[case] {msg}", flush=True)


class _DevBuf:
    """Caller-owned device memory (the test allocates; the entry point does not)."""

    def __init__(self, host: np.ndarray):
        self.hip = C.CDLL("libamdhip64.so")
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0
        assert self.hip.hipMemcpy(self.ptr, host.ctypes.data_as(C.c_void_p),
                                  C.c_size_t(self.nbytes), H2D) == 0

    def get(self, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr,
                                  C.c_size_t(self.nbytes), D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _solve_noalloc(ctx, A, B):
    """A: (nb, n, n), B: (nb, n, nrhs) complex -> X from the no-allocation entry."""
    from pssfss_amd.hip import lib
    nb, n, _ = A.shape
    nrhs = B.shape[2]
    Af = np.ascontiguousarray(np.transpose(A, (0, 2, 1))).view(np.float64)   # column-major
    Bf = np.ascontiguousarray(np.transpose(B, (0, 2, 1))).view(np.float64)
    dA, dB = _DevBuf(Af), _DevBuf(Bf)
    try:
        dp = C.POINTER(C.c_double)
        f = lib().pssfss_lu_solve_batch_noalloc
        f.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_int, dp, dp]
        f.restype = C.c_int
        ctx.check(f(ctx.h, n, nrhs, nb, C.cast(dA.ptr, dp), C.cast(dB.ptr, dp)))
        X = dB.get(Bf)
    finally:
        dA.free()
        dB.free()
    return np.transpose(X.view(np.complex128).reshape(nb, nrhs, n), (0, 2, 1))


def _system(rng, nb, n, nrhs):
    A = rng.standard_normal((nb, n, n)) + 1j * rng.standard_normal((nb, n, n))
    A += np.eye(n) * (2 * math.sqrt(n))
    B = rng.standard_normal((nb, n, nrhs)) + 1j * rng.standard_normal((nb, n, nrhs))
    return A, B


def test_noalloc_matches_allocating_solve(ctx):
    """Two calls in a row on the shared scratch, the second smaller: each solution
    equals pssfss_lu_solve_batch's FP64 solution to 1e-12 relative (the bar of
    the LU parity tests against LAPACK), and LAPACK's too."""
    rng = np.random.default_rng(11)
    for (nb, n, nrhs) in ((3, 700, 4), (2, 130, 2)):
        _say(f"no-allocation solve nbatch={nb} n={n} nrhs={nrhs}")
        A, B = _system(rng, nb, n, nrhs)
        X = _solve_noalloc(ctx, A, B)
        Xref, _, _ = ctx.lu_solve(A, B, "fp64")
        for b in range(nb):
            e1 = np.linalg.norm(X[b] - Xref[b]) / np.linalg.norm(Xref[b])
            e2 = np.linalg.norm(X[b] - np.linalg.solve(A[b], B[b])) / np.linalg.norm(Xref[b])
            print(f"  system {b}: vs allocating entry {e1:.2e}, vs LAPACK {e2:.2e}")
            assert e1 < 1e-12 and e2 < 1e-12


def test_noalloc_rejects_oversize(ctx):
    """A call beyond the fixed scratch is an argument error, checked on the host
    before any device work."""
    _say("no-allocation solve beyond the scratch limit")
    from pssfss_amd.hip import lib
    f = lib().pssfss_lu_solve_batch_noalloc
    dp = C.POINTER(C.c_double)
    f.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_int, dp, dp]
    f.restype = C.c_int
    dummy = np.zeros(2)
    st = f(ctx.h, 30000, 4, 1, dummy.ctypes.data_as(dp), dummy.ctypes.data_as(dp))
    assert st == -1
    assert "scratch" in lib().pssfss_last_error(ctx.h).decode()
