"""Test harness of the adjoint gradient: the displacement-mismatch objective of the
adjoint block, Newton continuation solves for finite differences, and the
torch.autograd.Function that exposes the HIP adjoint to torch (test harness only).

The objective is J(u) = sum over the observed top-surface nodes of |u - u_target|^2, with
u_target the forward solution for a second, independently seeded shear-modulus field.  The
parameters are log mu_e with K_e = 10 mu_e, so dJ/dlog mu_e = mu_e (10 dJ/dK_e + dJ/dmu_e).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import neon_amd
from neon_amd import problems
from oracle import oracle as orc

BULK_OVER_SHEAR = 10.0


def params_for(log_mu: np.ndarray) -> np.ndarray:
    mu = np.exp(log_mu)
    p = np.zeros((len(mu), 4))
    p[:, 0] = BULK_OVER_SHEAR * mu
    p[:, 1] = mu
    return p


def prescribed(p) -> np.ndarray:
    """Displacement holding only the Dirichlet values at the end of the load history."""
    u = np.zeros(p.n_dofs)
    for b in p.boundaries:
        u[np.asarray(b["dofs"])] = b["values"][-1]
    return u


def observed_dofs(p) -> np.ndarray:
    obs = np.asarray(p.extra["observed_nodes"])
    return (3 * obs[:, None] + np.arange(3)[None, :]).reshape(-1)


def objective(u, dofs, target):
    d = u[dofs] - target[dofs]
    return float(d @ d)


def objective_gradient(u, dofs, target):
    g = np.zeros_like(u)
    g[dofs] = 2.0 * (u[dofs] - target[dofs])
    return g


def log_mu_gradient(grad4: np.ndarray, log_mu: np.ndarray) -> np.ndarray:
    return np.exp(log_mu) * (BULK_OVER_SHEAR * grad4[:, 0] + grad4[:, 1])


# ------------------------------------------------------------------ GPU (HIP) side
def gpu_newton(mesh: neon_amd.DeviceMesh, u0: np.ndarray, fixed: np.ndarray, iterations: int = 3,
               tol: float = 1e-13) -> np.ndarray:
    """Full Newton from u0 (which carries the Dirichlet values) on the device primitives:
    K du = -f_int on the free dofs, with the consistent tangent."""
    u = u0.copy()
    for _ in range(iterations):
        mesh.set_displacement(u)
        mesh.update_internal_variables()
        mesh.assemble(neon_amd.PRECISION_FP64)
        _, f = mesh.download(stiffness=False)
        b = -f
        mesh.enforce_dirichlet(fixed, b)
        du, _, _ = mesh.solve(b, tol=tol, max_iterations=50000)
        u += du
    return u


def gpu_forward(p, log_mu, ctx, u_start=None, polish=2):
    """Converged displacement for log_mu: one load step of the Newton driver, then
    `polish` extra Newton iterations at a tight Krylov tolerance."""
    mesh = neon_amd.DeviceMesh(p.X, p.conn, params_for(log_mu), model=neon_amd.MODEL_NEOHOOKE,
                               tangent=neon_amd.TANGENT_CONSISTENT, context=ctx)
    fixed = p.fixed_dofs()
    if u_start is None:
        sim = dict(p.simulation)
        sm = neon_amd.static_matrix(mesh, sim, p.boundaries)
        u = sm.solve()
    else:
        u = u_start
    u = gpu_newton(mesh, u, fixed, iterations=polish)
    mesh.set_displacement(u)
    return mesh, u


def gpu_fd(p, log_mu, u0, direction, target, ctx, h=1e-6):
    """Central difference of J along `direction` in log mu, each side solved by Newton
    continuation from the converged u0."""
    fixed = p.fixed_dofs()
    dofs = observed_dofs(p)
    vals = []
    for sgn in (1.0, -1.0):
        mesh = neon_amd.DeviceMesh(p.X, p.conn, params_for(log_mu + sgn * h * direction),
                                   model=neon_amd.MODEL_NEOHOOKE, tangent=neon_amd.TANGENT_CONSISTENT, context=ctx)
        u = gpu_newton(mesh, u0, fixed, iterations=3)
        vals.append(objective(u, dofs, target))
    return (vals[0] - vals[1]) / (2.0 * h)


# ------------------------------------------------------------------ CPU oracle side
def oracle_newton(p, log_mu, u0, iterations=4):
    """Newton from u0 with the oracle's consistent tangent and a direct solve."""
    m = orc.Model(p.X, p.conn, params_for(log_mu), model=orc.MODEL_NEOHOOKE, tangent=orc.TANGENT_CONSISTENT)
    rp, ci = m.pattern()
    fixed = p.fixed_dofs()
    u = u0.copy()
    for _ in range(iterations):
        assert m.update(u) == orc.OK
        val, f = m.assemble(rp, ci)
        b = -f
        orc.enforce_dirichlet(rp, ci, val, b, fixed)
        u = u + spla.spsolve(sp.csr_matrix((val, ci, rp)).tocsc(), b)
    assert m.update(u) == orc.OK
    return m, u, rp, ci


def oracle_gradient(p, log_mu, u, dJdu):
    """Reference adjoint gradient: K lambda = dJ/du by a direct solve on the oracle's
    consistent tangent, then the oracle's per-element -lambda . df/dp."""
    m, u, rp, ci = oracle_newton(p, log_mu, u, iterations=1)
    val, _ = m.assemble(rp, ci)
    b = dJdu.copy()
    orc.enforce_dirichlet(rp, ci, val, b, p.fixed_dofs())
    lam = spla.spsolve(sp.csr_matrix((val, ci, rp)).tocsc(), b)
    return m.parameter_gradient(lam), lam


def oracle_fd(p, log_mu, u0, direction, target, h=1e-6):
    dofs = observed_dofs(p)
    vals = []
    for sgn in (1.0, -1.0):
        _, u, _, _ = oracle_newton(p, log_mu + sgn * h * direction, u0, iterations=3)
        vals.append(objective(u, dofs, target))
    return (vals[0] - vals[1]) / (2.0 * h)


def directions(n_elem, count):
    """Seeded random directions in log mu (unit RMS entries)."""
    return problems.normal(problems.SEED_DIRECTIONS, count * n_elem).reshape(count, n_elem)


# ------------------------------------------------------------------ torch wrapper
def make_autograd_objective(p, target, ctx):
    """J(log mu) as a torch.autograd.Function whose backward is the HIP adjoint."""
    import torch

    class AdjointObjective(torch.autograd.Function):
        @staticmethod
        def forward(fctx, log_mu):
            lm = log_mu.detach().cpu().numpy().astype(np.float64)
            mesh, u = gpu_forward(p, lm, ctx)
            dofs = observed_dofs(p)
            grad4 = mesh.adjoint_gradient(objective_gradient(u, dofs, target), p.fixed_dofs())
            fctx.save_for_backward(torch.from_numpy(log_mu_gradient(grad4, lm)))
            return torch.tensor(objective(u, dofs, target), dtype=torch.float64)

        @staticmethod
        def backward(fctx, g):
            (grad,) = fctx.saved_tensors
            return g * grad

    return AdjointObjective.apply
