"""The allocation-free CSR solve (neon_krylov_solve_csr_noalloc: device inputs and outputs, its
work vectors in the library's per-GPU static scratch) against the allocating entry point
neon_krylov_solve_csr on the same systems, through the C ABI: two calls in a row, the second on
a smaller system (the scratch keeps the first call's contents), for CG on an SPD matrix and
BiCGSTAB on an unsymmetric one.  Both entry points run the same kernels in the same order, so
iteration counts agree and the solutions agree to roundoff (expected bit for bit).  Also: a
system past the scratch's row limit and malformed CSR arrays are refused with
NEON_BAD_ARGUMENT before any solver kernel runs."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import neon_amd
from neon_amd import _capi

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 18


class _Device:
    """Device buffers through the HIP runtime that libneon_hip.so itself uses (its hipMalloc /
    hipMemcpy / hipFree, looked up through the library's handle), so no second runtime is
    initialised in the process."""

    def __init__(self):
        lib = _capi.lib()
        self.malloc, self.memcpy, self.free = lib.hipMalloc, lib.hipMemcpy, lib.hipFree
        self.malloc.restype = self.memcpy.restype = self.free.restype = C.c_int
        self.malloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.free.argtypes = [C.c_void_p]
        self.held = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.malloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.held.append(p)
        assert self.memcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.memcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def release(self):
        for p in self.held:
            self.free(p)
        self.held = []


def _noalloc():
    return _capi.lib().neon_krylov_solve_csr_noalloc


def _system(m, unsymmetric, seed):
    """7-point operator on an m^3 grid: -Laplacian + a random positive diagonal (SPD), plus a
    first-difference convection term along x when unsymmetric."""
    rng = np.random.default_rng(seed)
    one = sp.identity(m, format="csr")
    d2 = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(m, m), format="csr")
    L = sp.kron(sp.kron(d2, one), one) + sp.kron(sp.kron(one, d2), one) + sp.kron(sp.kron(one, one), d2)
    A = L + sp.diags(0.1 + rng.random(m ** 3))
    if unsymmetric:
        d1 = sp.diags([-1.0, 1.0], [-1, 0], shape=(m, m), format="csr")
        A = A + 0.4 * sp.kron(sp.kron(one, one), d1)
    A = sp.csr_matrix(A)
    A.sort_indices()
    b = rng.standard_normal(m ** 3)
    return A, b


def _solve_host(ctx, A, b, solver, criterion, tol):
    n = A.shape[0]
    rp, ci = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    v = A.data.astype(np.float64)
    x = np.zeros(n)
    its, err = C.c_int32(0), C.c_double(0)
    code = _capi.lib().neon_krylov_solve_csr(ctx.handle, solver, criterion, _capi.PRECISION_FP64, n, int(rp[-1]),
                                             _capi.i32(rp), _capi.i32(ci), _capi.f64(v), _capi.f64(b), _capi.f64(x),
                                             tol, 20000, C.byref(its), C.byref(err))
    ctx.check(code)
    return x, its.value, err.value


def _solve_noalloc(ctx, A, b, solver, criterion, tol):
    n = A.shape[0]
    dev = _Device()
    try:
        rp = dev.put(A.indptr.astype(np.int32))
        ci = dev.put(A.indices.astype(np.int32))
        v = dev.put(A.data.astype(np.float64))
        bd = dev.put(b)
        xd = dev.put(np.full(n, np.nan))  # overwritten by the solve
        its, err = C.c_int32(0), C.c_double(0)
        code = _noalloc()(ctx.handle, solver, criterion, n, int(A.indptr[-1]), rp, ci, v, bd, xd, tol, 20000,
                          C.byref(its), C.byref(err))
        ctx.check(code)
        return dev.get(xd, b), its.value, err.value
    finally:
        dev.release()


@pytest.mark.parametrize("solver,criterion,unsym", [
    (_capi.SOLVER_CG, _capi.STOP_RELATIVE, False),
    (_capi.SOLVER_CG, _capi.STOP_ABSOLUTE_PRECONDITIONED, False),
    (_capi.SOLVER_BICGSTAB, _capi.STOP_RELATIVE, True),
    (_capi.SOLVER_BICGSTAB, _capi.STOP_ABSOLUTE_PRECONDITIONED, True),
])
def test_noalloc_solve_equals_allocating_solve(ctx, solver, criterion, unsym):
    tol = 1e-10 if criterion == _capi.STOP_RELATIVE else 1e-12
    # 64^3 = 262,144 rows (the scratch's limit), then 20^3 rows in the same scratch
    for m, seed in ((64, 1), (20, 2)):
        A, b = _system(m, unsym, seed)
        assert A.shape[0] <= MAX_ROWS
        x0, its0, err0 = _solve_host(ctx, A, b, solver, criterion, tol)
        x1, its1, err1 = _solve_noalloc(ctx, A, b, solver, criterion, tol)
        diff = np.abs(x1 - x0).max() / np.abs(x0).max()
        print(f"m={m} solver={solver} criterion={criterion}: {its1} vs {its0} iterations, "
              f"error {err1:.3e} vs {err0:.3e}, max relative difference {diff:.3e}")
        assert its1 == its0 and its1 > 0
        assert diff <= 1e-13, diff
        # and the solve is a solve: the true residual is small.  Not for BiCGSTAB under the
        # CUDA rule, whose loop (biconjugate_gradient_stabilised_cuda.cpp:66, 181) tests
        # sqrt(r0 . r), which turns NaN when r0 . r < 0 and so ends the loop early (as here),
        # reporting a NaN error: both entry points reproduce that
        if not (solver == _capi.SOLVER_BICGSTAB and criterion == _capi.STOP_ABSOLUTE_PRECONDITIONED):
            assert np.linalg.norm(b - A @ x1) <= 1e-6 * np.linalg.norm(b)


def test_noalloc_refuses_what_does_not_fit(ctx):
    fn = _noalloc()
    its, err = C.c_int32(0), C.c_double(0)
    dev = _Device()
    try:
        # one row past the limit: refused on the host, nothing launched
        n = MAX_ROWS + 1
        ident = np.arange(n + 1, dtype=np.int32)
        rp, ci = dev.put(ident), dev.put(ident[:n])
        v, bd, xd = dev.put(np.ones(n)), dev.put(np.ones(n)), dev.put(np.zeros(n))
        code = fn(ctx.handle, _capi.SOLVER_CG, _capi.STOP_RELATIVE, n, n, rp, ci, v, bd, xd, 1e-10, 100,
                  C.byref(its), C.byref(err))
        assert code == _capi.NEON_BAD_ARGUMENT
        assert "at most 262144 rows" in _capi.lib().neon_context_last_error(ctx.handle).decode()
        # a column index out of range (in-bounds arrays, bad contents): refused by the check kernel
        n = 1000
        ident = np.arange(n + 1, dtype=np.int32)
        bad = ident[:n].copy()
        bad[17] = n
        b = np.linspace(1.0, 2.0, n)
        rp, ci_bad, ci = dev.put(ident), dev.put(bad), dev.put(ident[:n])
        v, bd, xd = dev.put(np.ones(n)), dev.put(b), dev.put(np.zeros(n))
        code = fn(ctx.handle, _capi.SOLVER_CG, _capi.STOP_RELATIVE, n, n, rp, ci_bad, v, bd, xd, 1e-10, 100,
                  C.byref(its), C.byref(err))
        assert code == _capi.NEON_BAD_ARGUMENT
        assert "malformed CSR" in _capi.lib().neon_context_last_error(ctx.handle).decode()
        # the corrected system solves: x = b for the identity (one CG step, alpha = 1 exactly)
        code = fn(ctx.handle, _capi.SOLVER_CG, _capi.STOP_RELATIVE, n, n, rp, ci, v, bd, xd, 1e-10, 100,
                  C.byref(its), C.byref(err))
        ctx.check(code)
        assert np.array_equal(dev.get(xd, b), b)
    finally:
        dev.release()
