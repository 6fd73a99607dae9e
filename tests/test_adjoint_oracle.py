"""Pinning of the oracle's adjoint gradient on the CPU (no GPU needed).

* Euler's theorem: the neo-Hooke stress is homogeneous of degree 1 in (bulk, shear), so for
  any lambda, sum_k p_e,k grad_e,k = -lambda_e . f_e with f_e the oracle's element force.
* Central finite differences (step 1e-6 in log mu) of the converged objective along seeded
  random directions, every solve a Newton continuation with a direct linear solve.
"""
import numpy as np
import pytest

from neon_amd import problems
from oracle import oracle as orc

import adjoint_harness as ah


def _setup(n):
    p = problems.adjoint_block(n, stretch=0.05, n_obs=64)
    log_mu = np.log(p.extra["mu"])
    u0 = ah.prescribed(p)
    _, target, _, _ = ah.oracle_newton(p, np.log(p.extra["mu_target"]), u0, iterations=8)
    _, u, _, _ = ah.oracle_newton(p, log_mu, u0, iterations=8)
    return p, log_mu, u, target


def test_parameter_gradient_euler_identity():
    p, log_mu, u, _ = _setup(3)
    m, u, _, _ = ah.oracle_newton(p, log_mu, u, iterations=1)
    lam = np.random.default_rng(0).normal(size=p.n_dofs)
    g = m.parameter_gradient(lam)
    params = ah.params_for(log_mu)
    conn = p.conn.reshape(-1, 8)
    for e in range(p.n_elem):
        _, fe = m.element(e)
        le = lam[(3 * conn[e][:, None] + np.arange(3)[None, :]).reshape(-1)]
        assert params[e] @ g[e] == pytest.approx(-le @ fe, rel=1e-12, abs=1e-14 * np.abs(fe).max())
    assert (g[:, 2:] == 0).all()


def test_oracle_adjoint_matches_finite_differences():
    p, log_mu, u, target = _setup(4)
    dofs = ah.observed_dofs(p)
    grad4, _ = ah.oracle_gradient(p, log_mu, u, ah.objective_gradient(u, dofs, target))
    g = ah.log_mu_gradient(grad4, log_mu)
    for d in ah.directions(p.n_elem, 4):
        fd = ah.oracle_fd(p, log_mu, u, d, target)
        adj = g @ d
        assert abs(fd - adj) <= 1e-6 * abs(adj), (fd, adj)


# ------------------------------------------------------------- J2 load-history adjoint
def _j2_case():
    p = problems.notched_plate_j2(8, 4, 2, steps=5, hole_cells=1.0, residual_tolerance=1e-12,
                                  displacement_tolerance=1e-12, norm_mode=1, solver_tolerance=1e-14,
                                  max_iterations=100000)
    w = np.random.default_rng(21).normal(size=p.n_dofs)
    w[p.fixed_dofs()] = 0.0
    return p, w


def _j2_forward(p, params, record=False):
    m = orc.Model(p.X, p.conn, params, model=orc.MODEL_J2, tangent=orc.TANGENT_CONSISTENT)
    code, u, recs = m.newton(orc.newton_options(p.simulation, record_trajectory=record), p.boundaries)
    assert code == orc.OK
    return m, u, recs


def test_oracle_j2_adjoint_matches_finite_differences():
    # objective J = w . u_N (a fixed linear functional of the final displacement); gradient
    # with respect to every element's yield stress, hardening, bulk and shear moduli
    p, w = _j2_case()
    params = p.params.copy()
    m, u, recs = _j2_forward(p, params, record=True)
    assert m.trajectory_steps() == 5
    q = m.qp_state()
    assert (q["ebar"] > 0).any() and (q["ebar"] == 0).any()  # part of the plate yielded
    g = m.adjoint_j2(w, p.fixed_dofs())
    rng = np.random.default_rng(5)
    for k, rel in ((2, 1e-6), (3, 1e-5), (0, 1e-6), (1, 1e-6)):
        d = rng.normal(size=p.n_elem)
        h = rel * np.abs(params[:, k]).mean()
        vals = []
        for sgn in (1.0, -1.0):
            pp = params.copy()
            pp[:, k] += sgn * h * d
            _, up, _ = _j2_forward(p, pp)
            vals.append(w @ up)
        fd = (vals[0] - vals[1]) / (2 * h)
        adj = g[:, k] @ d
        assert abs(fd - adj) <= 1e-6 * abs(adj) + 1e-9 * np.abs(g[:, k]).sum() * np.abs(d).max(), (k, fd, adj)
