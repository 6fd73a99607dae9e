"""The reference's example inputs and nonfollower loads through ingest and the device Newton
(neon_newton_solve / neon_newton_solve_loaded), against the CPU oracle on the same problem and
against the closed forms of test_ingest.py."""
import os

import numpy as np
import pytest

import neon_amd
from neon_amd import ingest
from oracle import oracle as orc

import neon_files
import test_ingest as ti

pytestmark = pytest.mark.gpu


def gpu_solve(p, ctx, precision="fp64"):
    mesh = neon_amd.DeviceMesh(p.X, p.conn, p.params, model=p.model, tangent=neon_amd.TANGENT_REFERENCE, context=ctx)
    sim = dict(p.simulation)
    sim["linear_solver"] = dict(sim["linear_solver"], backend={"type": "hip", "precision": precision})
    sm = neon_amd.static_matrix(mesh, sim, p.boundaries, loads=p.loads)
    return sm.solve(), sm.records


def _same_newton(recs, recs0):
    steps = lambda rr: [(round(r["step_time"], 12), r["iteration"]) for r in rr]  # noqa: E731
    assert steps(recs) == steps(recs0)
    assert all(r["status"] == 0 for r in recs)


def test_neohooke_cube_example(ctx):
    p = ingest.read(ti.NEOHOOKE_CUBE)
    u, recs = gpu_solve(p, ctx)
    code, u0, recs0, _ = ti.oracle_solve(p)
    assert code == orc.OK
    _same_newton(recs, recs0)
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-10 * np.abs(u0).max())
    U = u.reshape(-1, 3)
    assert U.max() == pytest.approx(0.5, rel=1e-9)
    axis = np.nonzero((p.X[:, 0] == 0) & (p.X[:, 1] == 0))[0]
    slope, icpt = np.polyfit(p.X[axis, 2], U[axis, 2], 1)
    assert slope == pytest.approx(0.5, rel=1e-7) and abs(icpt) < 1e-7


def test_j2_cube_example(ctx):
    p = ingest.read(ti.J2_CUBE)
    u, recs = gpu_solve(p, ctx)
    code, u0, recs0, _ = ti.oracle_solve(p)
    _same_newton(recs, recs0)
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-10 * np.abs(u0).max())
    K, G, sy, H = p.params[0]
    E = 9 * K * G / (3 * K + G)
    nu = (3 * K - 2 * G) / (2 * (3 * K + G))
    sig = (sy + H * 1e-3) * E / (E + H)
    elat = -nu * sig / E - 0.5 * (1e-3 - sig / E)
    U = u.reshape(-1, 3)
    np.testing.assert_allclose(U[:, 2], 1e-3 * p.X[:, 2], rtol=0, atol=1e-13)
    np.testing.assert_allclose(U[:, 0], elat * (p.X[:, 0] - 1.0), rtol=0, atol=1e-13)


@pytest.mark.parametrize("kind, value, t", [("traction", 0.4, 0.4), ("pressure", -0.4, 0.4), ("pressure", 0.3, -0.3)])
def test_uniaxial_loads(ctx, tmp_path, kind, value, t):
    p = ti._traction_problem(tmp_path, kind, value)
    u, recs = gpu_solve(p, ctx)
    code, u0, recs0, _ = ti.oracle_solve(p)
    _same_newton(recs, recs0)
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-10 * np.abs(u0).max())
    lz, ll = ti.uniaxial_neohooke_traction(5.0, 1.0, t)
    U = u.reshape(-1, 3)
    np.testing.assert_allclose(U[:, 2], (lz - 1.0) * p.X[:, 2], rtol=0, atol=1e-9)
    np.testing.assert_allclose(U[:, 0], (ll - 1.0) * p.X[:, 0], rtol=0, atol=1e-9)


def test_body_and_nodal_force(ctx, tmp_path):
    bcs = neon_files.rollers([
        {"name": "cube", "type": "body_force", "time": [0.0, 1.0], "z": [0.0, -0.3]},
        {"name": "z1", "type": "nodal_force", "time": [0.0, 1.0], "z": [0.0, 0.01]},
    ])
    p = ingest.read(neon_files.write_cube_simulation(tmp_path, 3, ti.NEOHOOKE, {"name": "neohooke"}, bcs))
    u, recs = gpu_solve(p, ctx)
    code, u0, recs0, _ = ti.oracle_solve(p)
    _same_newton(recs, recs0)
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-10 * np.abs(u0).max())


def test_loads_mixed_precision(ctx, tmp_path):
    # the mixed path under a traction: same converged displacement as the FP64 oracle
    p = ti._traction_problem(tmp_path, "traction", 0.4, n=4)
    u, recs = gpu_solve(p, ctx, precision="mixed")
    code, u0, recs0, _ = ti.oracle_solve(p)
    assert [round(r["step_time"], 12) for r in recs if r["iteration"] == 0] == \
        [round(r["step_time"], 12) for r in recs0 if r["iteration"] == 0]
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-9 * np.abs(u0).max())


def test_load_history_must_cover_the_steps(ctx, tmp_path):
    # a traction history ending before the step time: neon interpolates past its table there;
    # the device driver stops with an error naming the load history
    bcs = neon_files.rollers([
        {"name": "z1", "type": "displacement", "time": [0.0, 1.0], "z": [0.0, 0.01]},
        {"name": "x1", "type": "traction", "time": [0.0, 0.5], "x": [0.0, 0.1]},
    ])
    p = ingest.read(neon_files.write_cube_simulation(tmp_path, 2, ti.NEOHOOKE, {"name": "neohooke"}, bcs))
    with pytest.raises(neon_amd.NeonError, match="load's time history"):
        gpu_solve(p, ctx)


def test_generated_sinusoidal_history(ctx, tmp_path):
    """A generated (sinusoidal) boundary history drives the device Newton solve through the
    same steps and iterations as the oracle to the same displacement."""
    p = ti.sinusoid_problem(tmp_path, n=3)
    u, recs = gpu_solve(p, ctx)
    code, u0, recs0, _ = ti.oracle_solve(p)
    assert code == orc.OK
    _same_newton(recs, recs0)
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-10 * np.abs(u0).max())


def test_viscoplastic_damage_cube_example(ctx, tmp_path):
    # the reference's viscoplastic damage cube through ingest, the device Newton driver with
    # BiCGSTAB (the model's tangent is unsymmetric) and the output writer with the example's
    # fields: the oracle's Newton sequence, u and damage, and the damage fields in the files
    from neon_amd import output

    p = ingest.read(ti.DAMAGE_CUBE)
    u, sm, writer = output.solve_with_output(p, ctx, root=str(tmp_path))
    assert sm.options.solver == neon_amd.SOLVER_BICGSTAB
    code, u0, recs0, m = ti.oracle_solve(p)
    assert code == orc.OK
    _same_newton(sm.records, recs0)
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-9 * np.abs(u0).max())
    d, d0 = sm.mesh.damage_state()["damage"], m.damage_state()["damage"]
    assert (d0 > 0).all()
    np.testing.assert_allclose(d, d0, rtol=1e-8, atol=0)
    # one file per converged load step (10) after the initial write
    assert len(writer.written) >= 10
    last = open(os.path.join(str(tmp_path), writer.written[-1])).read()
    for name in ("damage", "energy_release_rate", "back_stress", "kinematic_hardening", "von_mises_stress"):
        assert f'Name="{name}"' in last, name
