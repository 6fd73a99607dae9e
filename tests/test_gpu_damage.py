"""small_strain_J2_plasticity_damage on the device against the CPU oracle (through the C ABI).

The device runs the same algorithm as the oracle (the 16-equation return map with complete
pivoting, neon's Mandel-form tangent, history updated in place at every call), so per-QP
states agree to rounding: stresses, tangents, plastic strain, back stress and damage to 1e-9 of
their size over a sequence of updates with and without a time step and a revert; the
unsymmetric tangent assembles into K (row-scaled 1e-10); Newton runs of the reference's cube
example (one element) and of a 3^3 block with BiCGSTAB take the oracle's steps and land on its
displacement to 1e-9."""
import numpy as np
import pytest

import neon_amd
from neon_amd import problems
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

CONST = problems.DAMAGE_CONSTANTS


def _pair(p, ctx):
    mg = neon_amd.DeviceMesh(p.X, p.conn, p.params, model=neon_amd.MODEL_J2_DAMAGE, context=ctx, damage=CONST)
    mr = orc.Model(p.X, p.conn, p.params, model=orc.MODEL_J2_DAMAGE)
    mr.set_damage(CONST)
    return mg, mr


def _close(a, b, tol, what):
    scale = max(np.abs(b).max(), 1e-30)
    err = np.abs(a - b).max() / scale
    assert err <= tol, (what, err)


def test_traits_and_names(ctx):
    p = problems.damage_cube(1)
    mg, _ = _pair(p, ctx)
    cm = mg.constitutive("steel")
    assert not cm.is_finite_deformation() and not cm.is_symmetric() and not mg.is_symmetric()
    assert {"damage", "back_stress", "energy_release_rate", "kinematic_hardening"} <= cm.variable_names()


def test_qp_state_sequence_parity(ctx):
    p = problems.damage_cube(2)
    mg, mr = _pair(p, ctx)
    X = p.X
    for k, (ez, dt, commit) in enumerate([(0.0008, 1e-3, True), (0.0012, 1e-3, False), (0.0012, 0.0, True),
                                          (0.0016, 1e-3, True), (0.0020, 1e-3, None)]):
        u = np.zeros_like(X)
        u[:, 2] = ez * X[:, 2]
        u[:, 0] = -0.3 * ez * X[:, 0] + 1e-5 * np.sin(3 * X[:, 1])
        u = u.reshape(-1)
        mg.set_displacement(u)
        mg.update_internal_variables(dt)
        assert mr.update(u, dt) == orc.OK
        sg, sr = mg.qp_state(), mr.qp_state()
        dg, dr = mg.damage_state(), mr.damage_state()
        for key in ("sigma", "D", "eps_p", "ebar", "vm"):
            _close(sg[key], sr[key], 1e-9, (k, key))
        for key in ("back_stress", "damage", "energy"):
            _close(dg[key], dr[key], 1e-9, (k, key))
        if commit is not None:
            mg.save_internal_variables(commit)
            mr.save(commit)
    assert (mr.damage_state()["damage"] > 0).all()
    # the assembled unsymmetric tangent
    mg.assemble(neon_amd.PRECISION_FP64)
    val, f = mg.download()
    rp, ci = mr.pattern()
    val0, f0 = mr.assemble(rp, ci)
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    rowmax = np.maximum.reduceat(np.abs(val0), rp[:-1])
    assert (np.abs(val - val0) / rowmax[rows]).max() < 1e-10
    np.testing.assert_allclose(f, f0, rtol=0, atol=1e-10 * np.abs(f0).max())
    import scipy.sparse as sp
    A = sp.csr_matrix((val0, ci, rp))
    # genuinely unsymmetric: well above the roundoff a symmetric assembly would leave
    assert abs(A - A.T).max() > 1e-9 * abs(A).max()


@pytest.mark.parametrize("n,top", [(1, 0.018), (3, 0.016)])
def test_newton_cube_parity(ctx, n, top):
    # neon's viscoplastic damage cube example (one element, top face pulled to 1.8 % in 10 steps
    # of 1 ms) and a 3^3 block to 1.6 % (at 1.8 % its last tangent is so near singular that
    # BiCGSTAB's convergence follows the summation order, also between the oracle's thread
    # counts); rollers, BiCGSTAB (Eigen's rule, relative 1e-10)
    p = problems.damage_cube(n, top=top, residual_tolerance=1e-8, displacement_tolerance=1e-8, norm_mode=1)
    mg, mr = _pair(p, ctx)
    sm = neon_amd.static_matrix(mg, p.simulation, p.boundaries)
    assert sm.options.solver == neon_amd.SOLVER_BICGSTAB
    u = sm.solve()
    code, u0, recs0 = mr.newton(orc.newton_options(p.simulation, solver=1), p.boundaries)
    assert code == orc.OK
    steps = [r["step"] for r in sm.records]
    steps0 = [r["step"] for r in recs0]
    assert steps == steps0, (steps, steps0)
    err = np.abs(u - u0).max() / np.abs(u0).max()
    print(f"damage cube {n}: {len(sm.records)} Newton iterations, |u - u_oracle| / max|u| = {err:.2e}")
    assert err < 1e-9
    dg, dr = mg.damage_state(), mr.damage_state()
    _close(dg["damage"], dr["damage"], 1e-8, "damage")
    assert (dr["damage"] > 0).all()
