"""Adjoint gradient of the HIP path (neon_adjoint_gradient through the C ABI).

Bars: the gradient in log mu agrees with the CPU oracle's reference gradient (direct
solves) to 1e-8 relative in the 2-norm, and with FP64 central differences (step 1e-6 in
log mu, Newton continuation solves on the device) to 1e-6 relative along each of 32 seeded
random directions; the torch.autograd.Function wrapper returns the same gradient."""
import numpy as np
import pytest

import neon_amd
from neon_amd import problems

import adjoint_harness as ah

pytestmark = pytest.mark.gpu


def _setup(n, ctx, n_obs=4096):
    p = problems.adjoint_block(n, stretch=0.05, n_obs=n_obs)
    log_mu = np.log(p.extra["mu"])
    _, target = ah.gpu_forward(p, np.log(p.extra["mu_target"]), ctx)
    mesh, u = ah.gpu_forward(p, log_mu, ctx)
    return p, log_mu, mesh, u, target


def test_adjoint_matches_oracle_gradient(ctx):
    p, log_mu, mesh, u, target = _setup(8, ctx, n_obs=40)
    dofs = ah.observed_dofs(p)
    dJdu = ah.objective_gradient(u, dofs, target)
    g = ah.log_mu_gradient(mesh.adjoint_gradient(dJdu, p.fixed_dofs(), tol=1e-13), log_mu)
    g4_ref, _ = ah.oracle_gradient(p, log_mu, u, dJdu)
    g_ref = ah.log_mu_gradient(g4_ref, log_mu)
    err = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    assert err < 1e-8, err


def test_adjoint_matches_finite_differences(ctx):
    p, log_mu, mesh, u, target = _setup(16, ctx)
    dofs = ah.observed_dofs(p)
    grad4 = mesh.adjoint_gradient(ah.objective_gradient(u, dofs, target), p.fixed_dofs(), tol=1e-13)
    assert mesh.adjoint_iterations > 0
    assert (grad4[:, 2:] == 0).all()
    g = ah.log_mu_gradient(grad4, log_mu)
    worst = 0.0
    for d in ah.directions(p.n_elem, 32):
        fd = ah.gpu_fd(p, log_mu, u, d, target, ctx)
        adj = g @ d
        worst = max(worst, abs(fd - adj) / abs(adj))
    assert worst < 1e-6, worst


def test_adjoint_through_torch_autograd(ctx):
    torch = pytest.importorskip("torch")
    p, log_mu, mesh, u, target = _setup(6, ctx)
    dofs = ah.observed_dofs(p)
    g = ah.log_mu_gradient(mesh.adjoint_gradient(ah.objective_gradient(u, dofs, target), p.fixed_dofs(), tol=1e-13),
                           log_mu)
    f = ah.make_autograd_objective(p, target, ctx)
    x = torch.tensor(log_mu, dtype=torch.float64, requires_grad=True)
    J = f(x)
    J.backward()
    assert float(J) == pytest.approx(ah.objective(u, dofs, target), rel=1e-10)
    np.testing.assert_allclose(x.grad.numpy(), g, rtol=1e-9, atol=1e-12 * np.abs(g).max())


def test_adjoint_rejects_reference_tangent(ctx):
    p = problems.adjoint_block(3, n_obs=4)
    mesh = neon_amd.DeviceMesh(p.X, p.conn, p.params, context=ctx)
    with pytest.raises(neon_amd.NeonError):
        mesh.adjoint_gradient(np.zeros(p.n_dofs), p.fixed_dofs())


# ------------------------------------------------------------- J2 load-history adjoint
def _plate(nx, ny, nz, steps, hole):
    p = problems.notched_plate_j2(nx, ny, nz, steps=steps, hole_cells=hole, residual_tolerance=1e-12,
                                  displacement_tolerance=1e-12, norm_mode=1, solver_tolerance=1e-14,
                                  max_iterations=100000)
    p.simulation["record_trajectory"] = True
    w = np.random.default_rng(21).normal(size=p.n_dofs)
    w[p.fixed_dofs()] = 0.0
    return p, w


def _j2_gpu_forward(p, params, ctx):
    mesh = neon_amd.DeviceMesh(p.X, p.conn, params, model=neon_amd.MODEL_J2, tangent=neon_amd.TANGENT_CONSISTENT,
                               context=ctx)
    sm = neon_amd.static_matrix(mesh, p.simulation, p.boundaries)
    u = sm.solve()
    return mesh, u


def test_j2_adjoint_matches_oracle_gradient(ctx):
    from oracle import oracle as orc

    p, w = _plate(8, 4, 2, 5, 1.0)
    mesh, u = _j2_gpu_forward(p, p.params, ctx)
    before = mesh.qp_state()
    g = mesh.adjoint_gradient(w, p.fixed_dofs(), tol=1e-14)
    after = mesh.qp_state()
    # the converged final state is restored
    for key in ("sigma", "eps_p", "ebar"):
        np.testing.assert_array_equal(after[key], before[key])
    ref = orc.Model(p.X, p.conn, p.params, model=orc.MODEL_J2, tangent=orc.TANGENT_CONSISTENT)
    code, u0, _ = ref.newton(orc.newton_options(p.simulation, record_trajectory=True), p.boundaries)
    assert code == orc.OK
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-10 * np.abs(u0).max())
    g0 = ref.adjoint_j2(w, p.fixed_dofs())
    for k in range(4):
        err = np.linalg.norm(g[:, k] - g0[:, k]) / np.linalg.norm(g0[:, k])
        assert err < 1e-7, (k, err)


def test_j2_adjoint_matches_finite_differences(ctx):
    # notched plate, 20 load steps to 0.5 % strain, objective w . u_N; yield-stress gradient
    # against central differences (step 1e-6 relative) of full forward re-solves
    p, w = _plate(12, 6, 3, 20, 1.5)
    mesh, u = _j2_gpu_forward(p, p.params, ctx)
    assert (mesh.qp_state()["ebar"] > 0).any()
    g = mesh.adjoint_gradient(w, p.fixed_dofs(), tol=1e-14)
    rng = np.random.default_rng(8)
    worst = 0.0
    for _ in range(6):
        d = rng.normal(size=p.n_elem)
        h = 1e-6 * np.abs(p.params[:, 2]).mean()
        vals = []
        for sgn in (1.0, -1.0):
            pp = p.params.copy()
            pp[:, 2] += sgn * h * d
            _, up = _j2_gpu_forward(p, pp, ctx)
            vals.append(w @ up)
        fd = (vals[0] - vals[1]) / (2 * h)
        adj = g[:, 2] @ d
        worst = max(worst, abs(fd - adj) / abs(adj))
    assert worst < 1e-6, worst


def test_j2_adjoint_needs_a_recorded_history(ctx):
    p, w = _plate(4, 2, 2, 2, 0.0)
    p.simulation["record_trajectory"] = False
    mesh, u = _j2_gpu_forward(p, p.params, ctx)
    with pytest.raises(neon_amd.NeonError):
        mesh.adjoint_gradient(w, p.fixed_dofs())
