"""The ageing microsphere (gaussian_ageing_affine_microsphere with the BAND network) on the HIP
path, through the C ABI, against the CPU oracle (pinned in test_microsphere.py).

Tolerances as for the other models: network state, stress and tangent per QP to 1e-12 of their
maximum (the network's Runge-Kutta step and the secondary stress are the oracle's operations in
the same order, up to fused multiply-adds); Newton runs with equal step and iteration sequences
and u to 1e-10 of max |u|."""
import numpy as np
import pytest

import neon_amd
from oracle import oracle as orc

import test_gpu_models as tgm
import test_microsphere as tms

pytestmark = pytest.mark.gpu

PROBABILITIES = (1.0e-3, 2.0e-3)


def _params(n_elem, seed=11):
    rng = np.random.default_rng(seed)
    K = 3.0 * rng.uniform(0.8, 1.25, n_elem)
    G = rng.uniform(0.8, 1.25, n_elem)
    N = rng.uniform(20.0, 40.0, n_elem)
    cure = rng.uniform(40.0, 100.0, n_elem)
    return np.stack([K, G, N, (1.0 - cure / 100.0) * G], axis=1)


def _scale(z):
    """Per-column magnitude of a network state, at least 1e-3 of its group's (the five network
    scalars, the six secondary stress components), so that a column of zeros is compared
    against the stress level rather than bit for bit."""
    col = np.abs(z).max(axis=0)
    floor = np.concatenate([np.full(5, col[:5].max()), np.full(6, col[5:].max())])
    return np.maximum(col, 1e-3 * floor)


def _within(za, zr, tol):
    err = np.abs(za - zr) / _scale(zr)
    worst = np.unravel_index(np.argmax(err), err.shape)
    assert err.max() <= tol, f"network state {worst}: {za[worst]!r} vs {zr[worst]!r} (scaled error {err.max():.3e})"


def _compare(m_gpu, m_ref):
    _within(m_gpu.ageing_state(), m_ref.ageing_state(), 1e-12)
    sg, sr = m_gpu.qp_state(), m_ref.qp_state()
    for key in ("sigma", "D", "detF"):
        np.testing.assert_allclose(sg[key], sr[key], rtol=0, atol=1e-12 * np.abs(sr[key]).max(), err_msg=key)


@pytest.mark.parametrize("rule", [0, orc.SPHERE_BO33])
def test_state_history_and_assembly_parity(ctx, rule):
    """Three updates with commit and revert between them: the network advances only on a
    non-zero dt, the secondary stress builds from the committed value with the last non-zero
    step, and a revert restores the committed network."""
    print(f"ageing rule {rule:#x}", flush=True)
    X, conn = tgm._block()
    prm = _params(len(conn))
    model = orc.MODEL_MICROSPHERE_AGEING | rule
    m_gpu = neon_amd.DeviceMesh(X, conn, prm, model=model, context=ctx, ageing=PROBABILITIES)
    m_ref = orc.Model(X, conn, prm, model=model)
    m_ref.set_ageing(*PROBABILITIES)
    np.testing.assert_array_equal(m_gpu.ageing_state(), m_ref.ageing_state())
    for scale, dt, action in ((0.1, 1.0, "commit"), (0.15, 0.0, None), (0.2, 0.5, "revert"), (0.12, 2.0, "commit"),
                              (0.18, 0.0, None)):
        u = tgm._deformation(X, scale)
        m_gpu.set_displacement(u)
        m_gpu.update_internal_variables(dt)
        assert m_ref.update(u, dt) == orc.OK
        _compare(m_gpu, m_ref)
        if action:
            # the network is the path-dependent state; a revert also restores the oracle's
            # stress, which the device recomputes from u at the next update
            m_gpu.save_internal_variables(action == "commit")
            m_ref.save(action == "commit")
            _within(m_gpu.ageing_state(), m_ref.ageing_state(), 1e-12)
    rp, ci = m_ref.pattern()
    val0, f0 = m_ref.assemble(rp, ci)
    m_gpu.assemble(neon_amd.PRECISION_FP64)
    val, f = m_gpu.download()
    assert tgm._row_scaled(val, val0, rp) < 1e-12
    np.testing.assert_allclose(f, f0, rtol=0, atol=1e-12 * np.abs(f0).max())


def test_ageing_checks(ctx):
    X, conn = tgm._block(2)
    prm = _params(len(conn))
    with pytest.raises(neon_amd.NeonError):
        neon_amd.DeviceMesh(X, conn, prm, model=orc.MODEL_MICROSPHERE_AGEING, tangent=neon_amd.TANGENT_CONSISTENT,
                            context=ctx)
    m = neon_amd.DeviceMesh(X, conn, prm, model=orc.MODEL_MICROSPHERE_AGEING, context=ctx)
    with pytest.raises(neon_amd.NeonError):
        m.set_ageing(-1.0e-3, 0.0)
    other = neon_amd.DeviceMesh(X, conn, prm, model=orc.MODEL_MICROSPHERE_GAUSSIAN, context=ctx)
    with pytest.raises(neon_amd.NeonError):
        other.ageing_state()


@pytest.mark.parametrize("rule", ["BO21", "BO33"])
def test_newton_stretch_and_hold(ctx, tmp_path, rule):
    """A BAND cube stretched and held over four load steps: each step's first update carries
    its increment, so the device repeats the oracle's steps, iterations (the held steps move
    as the network ages) and displacement, and ends in the same network state."""
    p = tms.ageing_simulation(tmp_path, rule=rule)
    mesh = neon_amd.DeviceMesh(p.X, p.conn, p.params, model=p.model, tangent=neon_amd.TANGENT_REFERENCE,
                               context=ctx, ageing=p.ageing)
    sim = dict(p.simulation)
    sim["linear_solver"] = dict(sim["linear_solver"], backend={"type": "hip", "precision": "fp64"})
    sm = neon_amd.static_matrix(mesh, sim, p.boundaries, loads=p.loads)
    u = sm.solve()
    code, u0, recs0, m_ref = tms.oracle_ageing_solve(p)
    assert code == orc.OK
    steps = lambda rr: [(round(r["step_time"], 12), r["iteration"]) for r in rr]  # noqa: E731
    assert steps(sm.records) == steps(recs0)
    assert all(r["status"] == 0 for r in sm.records)
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-10 * np.abs(u0).max())
    _within(mesh.ageing_state(), m_ref.ageing_state(), 1e-10)
