"""Parity of the HIP path (through the C ABI) with the CPU FP64 oracle.

Tolerances (FP64 path): tangent entries |dK| <= 1e-12 * max|K_row| (row-scaled; K entries of
a regular mesh cancel to roundoff, so a per-entry relative bound is undefined there);
f_int and per-QP state to 1e-12 relative to their max; displacements to 1e-10 relative to
max|u|; Krylov and Newton iteration counts equal (Krylov allowed +-1 where a residual lands
within roundoff of the tolerance)."""
import numpy as np
import pytest
import scipy.sparse as sp

import neon_amd
from neon_amd import problems
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _deformed_state(p, scale, seed=0):
    rng = np.random.default_rng(seed)
    u = scale * rng.normal(size=p.n_dofs)
    return u


def _row_scaled_close(val_gpu, val_ref, rp, tol):
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    rowmax = np.maximum.reduceat(np.abs(val_ref), rp[:-1])
    err = np.abs(val_gpu - val_ref) / rowmax[rows]
    return err.max()


def _pair(p, ctx, tangent=neon_amd.TANGENT_REFERENCE):
    m_gpu = neon_amd.DeviceMesh(p.X, p.conn, p.params, model=p.model, tangent=tangent, context=ctx)
    m_ref = orc.Model(p.X, p.conn, p.params, model=p.model, tangent=tangent)
    return m_gpu, m_ref


@pytest.mark.parametrize("case", ["cube10", "block6"])
def test_pattern_matches_oracle(ctx, case):
    p = problems.cube_clamped(10) if case == "cube10" else problems.block_rollers(6)
    m_gpu, m_ref = _pair(p, ctx)
    rp, ci = m_gpu.pattern()
    rp0, ci0 = m_ref.pattern()
    np.testing.assert_array_equal(rp, rp0)
    np.testing.assert_array_equal(ci, ci0)


@pytest.mark.parametrize("case,tangent", [("cube10", 0), ("block6", 0), ("block6", 1)])
def test_state_and_assembly_parity_neohooke(ctx, case, tangent):
    p = problems.cube_clamped(10) if case == "cube10" else problems.block_rollers(6)
    m_gpu, m_ref = _pair(p, ctx, tangent)
    u = _deformed_state(p, 0.01, seed=1)
    m_gpu.set_displacement(u)
    m_gpu.update_internal_variables()
    assert m_ref.update(u) == orc.OK
    sg, sr = m_gpu.qp_state(), m_ref.qp_state()
    for key in ("sigma", "D", "detF"):
        ref = sr[key]
        np.testing.assert_allclose(sg[key], ref, rtol=0, atol=1e-12 * np.abs(ref).max(), err_msg=key)
    m_gpu.assemble(neon_amd.PRECISION_FP64)
    val, f = m_gpu.download()
    rp, ci = m_ref.pattern()
    val0, f0 = m_ref.assemble(rp, ci)
    assert _row_scaled_close(val, val0, rp, 0) < 1e-12
    np.testing.assert_allclose(f, f0, rtol=0, atol=1e-12 * np.abs(f0).max())


def test_state_and_assembly_parity_j2(ctx):
    p = problems.notched_plate_j2(12, 6, 3, hole_cells=1.5, jitter=0.1)
    m_gpu, m_ref = _pair(p, ctx)
    # uniaxial strain at the yield limit (q = 2 G eps ~ sigma_y) plus noise: part yields
    u = np.zeros(p.n_dofs)
    u[0::3] = 0.0016 * p.X[:, 0]
    u += _deformed_state(p, 2e-5, seed=2)
    for factor in (1.0, 1.05):  # the history accumulates across calls (in-place update)
        m_gpu.set_displacement(u * factor)
        m_gpu.update_internal_variables()
        assert m_ref.update(u * factor) == orc.OK
    sg, sr = m_gpu.qp_state(), m_ref.qp_state()
    assert (sr["ebar"] > 0).any() and (sr["ebar"] == 0).any()
    for key in ("sigma", "D", "eps_p", "ebar", "vm"):
        ref = sr[key]
        np.testing.assert_allclose(sg[key], ref, rtol=0, atol=1e-11 * max(np.abs(ref).max(), 1e-30), err_msg=key)
    m_gpu.assemble(neon_amd.PRECISION_FP64)
    val, f = m_gpu.download()
    rp, ci = m_ref.pattern()
    val0, f0 = m_ref.assemble(rp, ci)
    assert _row_scaled_close(val, val0, rp, 0) < 1e-12
    np.testing.assert_allclose(f, f0, rtol=0, atol=1e-12 * np.abs(f0).max())


def test_negative_jacobian_flag(ctx):
    p = problems.cube_clamped(2)
    m_gpu, _ = _pair(p, ctx)
    u = np.zeros(p.n_dofs)
    u[2::3] = -2.0 * p.X[:, 2]
    m_gpu.set_displacement(u)
    with pytest.raises(neon_amd.computational_error) as e:
        m_gpu.update_internal_variables()
    assert e.value.code == 3


def _system(p, ctx, scale=0.01):
    """Oracle tangent + rhs after Dirichlet, at a deformed state."""
    m_gpu, m_ref = _pair(p, ctx)
    u = _deformed_state(p, scale, seed=3)
    assert m_ref.update(u) == orc.OK
    rp, ci = m_ref.pattern()
    val, f = m_ref.assemble(rp, ci)
    b = -f + 1e-3 * np.random.default_rng(4).normal(size=p.n_dofs)
    fixed = p.fixed_dofs()
    orc.enforce_dirichlet(rp, ci, val, b, fixed)
    return m_gpu, m_ref, u, rp, ci, val, b, fixed


def test_dirichlet_parity(ctx):
    p = problems.block_rollers(6)
    m_gpu, m_ref = _pair(p, ctx)
    u = _deformed_state(p, 0.01, seed=5)
    m_gpu.set_displacement(u)
    m_gpu.update_internal_variables()
    m_gpu.assemble()
    assert m_ref.update(u) == orc.OK
    rp, ci = m_ref.pattern()
    val0, f0 = m_ref.assemble(rp, ci)
    b0 = -f0.copy()
    fixed = p.fixed_dofs()
    orc.enforce_dirichlet(rp, ci, val0, b0, fixed)
    b = -f0.copy()
    m_gpu.enforce_dirichlet(fixed, b)
    val, _ = m_gpu.download()
    assert _row_scaled_close(val, val0, rp, 0) < 1e-12
    assert (val[val0 == 0] == 0).all()
    np.testing.assert_array_equal(b, b0)


@pytest.mark.parametrize("criterion", [neon_amd.STOP_RELATIVE, neon_amd.STOP_ABSOLUTE_PRECONDITIONED])
def test_cg_host_csr_parity(ctx, criterion):
    p = problems.cube_clamped(10)
    _, _, _, rp, ci, val, b, _ = _system(p, ctx)
    tol = 1e-10 if criterion == neon_amd.STOP_RELATIVE else 1e-12
    ref = orc.cg_eigen if criterion == neon_amd.STOP_RELATIVE else orc.cg_cuda
    code, x0, its0, err0 = ref(rp, ci, val, b, tol, 20000)
    assert code == orc.OK
    s = neon_amd.conjugate_gradient_hip(tol, 20000, criterion=criterion, context=ctx)
    x = np.zeros_like(b)
    s.solve((rp, ci, val), x, b)
    assert abs(s.iterations - its0) <= 1
    np.testing.assert_allclose(x, x0, rtol=0, atol=1e-10 * np.abs(x0).max())


def test_cg_reference_tridiagonal_golden(ctx):
    # test/linear_solvers.cpp:15-49 through the drop-in class and the factory
    A = sp.csr_matrix(np.array([[2.0, -1.0, 0.0], [-1.0, 2.0, -1.0], [0.0, -1.0, 2.0]]))
    b = np.array([1.0, 0.5, 1.5])
    for data in ({"type": "iterative", "device": "gpu", "backend": {"type": "hip"}},
                 {"type": "iterative", "device": "gpu", "tolerance": 1e-8, "maximum_iterations": 100,
                  "backend": {"type": "hip"}}):
        s = neon_amd.make_linear_solver(data, True, context=ctx)
        x = b.copy()
        s.solve(A, x, b)
        np.testing.assert_allclose(x, [1.375, 1.75, 1.625], atol=1e-5)


def test_cg_iteration_limit_is_recoverable(ctx):
    p = problems.cube_clamped(6)
    _, _, _, rp, ci, val, b, _ = _system(p, ctx)
    s = neon_amd.conjugate_gradient_hip(1e-12, 3, context=ctx)
    x = np.zeros_like(b)
    with pytest.raises(neon_amd.computational_error) as e:
        s.solve((rp, ci, val), x, b)
    assert e.value.code == 1 and s.iterations == 3


def test_device_tangent_solve_parity(ctx):
    p = problems.block_rollers(8)
    m_gpu, m_ref = _pair(p, ctx)
    u = _deformed_state(p, 0.005, seed=6)
    m_gpu.set_displacement(u)
    m_gpu.update_internal_variables()
    m_gpu.assemble()
    assert m_ref.update(u) == orc.OK
    rp, ci = m_ref.pattern()
    val0, f0 = m_ref.assemble(rp, ci)
    b0 = -f0.copy()
    fixed = p.fixed_dofs()
    orc.enforce_dirichlet(rp, ci, val0, b0, fixed)
    b = -f0.copy()
    m_gpu.enforce_dirichlet(fixed, b)
    code, x0, its0, _ = orc.cg_eigen(rp, ci, val0, b0, 1e-10, 20000)
    x, its, err = m_gpu.solve(b, criterion=neon_amd.STOP_RELATIVE, tol=1e-10)
    assert abs(its - its0) <= 1
    np.testing.assert_allclose(x, x0, rtol=0, atol=1e-10 * np.abs(x0).max())


def _newton_pair(p, ctx, criterion=neon_amd.STOP_RELATIVE):
    m_gpu, m_ref = _pair(p, ctx)
    sm = neon_amd.static_matrix(m_gpu, dict(p.simulation, criterion=criterion), p.boundaries)
    u = sm.solve()
    code, u0, recs0 = m_ref.newton(orc.newton_options(p.simulation, criterion=criterion), p.boundaries)
    return u, sm.records, u0, recs0, code


def test_newton_parity_cube10(ctx):
    # the 10^3 clamped cube, 5 load steps, Jacobi-PCG 1e-10
    p = problems.cube_clamped(10)
    u, recs, u0, recs0, code = _newton_pair(p, ctx)
    assert code == orc.OK
    assert [(r["step"], r["iteration"]) for r in recs] == [(r["step"], r["iteration"]) for r in recs0]
    for r, r0 in zip(recs, recs0):
        assert abs(r["solver_iterations"] - r0["solver_iterations"]) <= 1
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-10 * np.abs(u0).max())


def test_newton_parity_block_relative_norms(ctx):
    # the jittered roller block recipe at 6^3, relative norms, 3 of 10 load steps
    p = problems.block_rollers(6, residual_tolerance=1e-8, displacement_tolerance=1e-8, norm_mode=1,
                               max_load_steps=0)
    p.simulation["time"]["period"] = 1.0
    u, recs, u0, recs0, code = _newton_pair(p, ctx)
    assert code == orc.OK
    assert [(r["step"], r["iteration"]) for r in recs] == [(r["step"], r["iteration"]) for r in recs0]
    for r, r0 in zip(recs, recs0):
        assert abs(r["solver_iterations"] - r0["solver_iterations"]) <= 1
        assert r["force_norm"] == pytest.approx(r0["force_norm"], rel=1e-6, abs=1e-14)
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-10 * np.abs(u0).max())


def test_newton_parity_j2_plate(ctx):
    p = problems.notched_plate_j2(12, 6, 3, steps=5, hole_cells=1.5, residual_tolerance=1e-8,
                                  displacement_tolerance=1e-8, norm_mode=1)
    u, recs, u0, recs0, code = _newton_pair(p, ctx)
    assert code == orc.OK
    assert [(r["step"], r["iteration"]) for r in recs] == [(r["step"], r["iteration"]) for r in recs0]
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-10 * np.abs(u0).max())
