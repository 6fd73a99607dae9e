"""Natural frequencies on the device (neon_natural_frequency, neon_mesh_mass) against the oracle's
K and M solved densely (scipy.linalg.eigh) and against the bar's closed form
(test_natural_frequency.py).

Tolerances: mass entries to 1e-14 of their maximum; eigenvalues to 1e-8 relative (the Ritz
residual tolerance is 1e-8, so the eigenvalues' error is of order 1e-16 times the spectrum's
spread over its gaps; the inner PCG solves go to 1e-14); mode shapes M-orthonormal to 1e-8
with relative residuals |K x - lambda M x| / |lambda M x| below 1e-6; repeated eigenvalues of a
symmetric mesh found with their multiplicity."""
import time

import numpy as np
import pytest
import scipy.linalg

import neon_amd
from neon_amd import ingest, problems
from oracle import oracle as orc

import test_natural_frequency as tnf

pytestmark = pytest.mark.gpu


def _device(X, conn, params, ctx, model=orc.MODEL_LINEAR_ELASTIC):
    return neon_amd.DeviceMesh(X, conn, params, model=model, context=ctx)


def _check_modes(lam, vecs, rp, ci, kv, mv):
    K, M = tnf._dense(rp, ci, kv), tnf._dense(rp, ci, mv)
    G = vecs.T @ M @ vecs
    np.testing.assert_allclose(G, np.eye(len(lam)), atol=1e-8)
    for i, l in enumerate(lam):
        Mx = M @ vecs[:, i]
        assert np.linalg.norm(K @ vecs[:, i] - l * Mx) <= 1e-6 * l * np.linalg.norm(Mx)


def test_mass_parity(ctx):
    X, conn, _ = problems.box_mesh(4, 3, 5, L=(1.0, 0.8, 1.3), jitter=0.1)
    params = np.tile([3.0, 1.0, 0, 0], (len(conn), 1))
    rho = np.random.default_rng(3).uniform(1.0, 3.0, len(conn))
    m = _device(X, conn, params, ctx)
    val = m.mass(rho)
    ref = orc.Model(X, conn, params, model=orc.MODEL_LINEAR_ELASTIC)
    rp, ci = ref.pattern()
    val0 = ref.assemble_mass(rp, ci, rho)
    np.testing.assert_allclose(val, val0, rtol=0, atol=1e-14 * np.abs(val0).max())


def test_bar_closed_form(ctx):
    N = 20
    X, conn, params, fixed = tnf.bar(N)
    m = _device(X, conn, params, ctx)
    fm = neon_amd.natural_frequency_matrix(m, {"type": "arpack", "eigenvalues": 8}, [{"dofs": fixed}], 1.0)
    lam, vecs = fm.solve()
    print("bar: solves", fm.solves, flush=True)
    np.testing.assert_allclose(lam, tnf.bar_eigenvalues(N, *params[0, :2])[:8], rtol=1e-8)
    rp, ci, kv, mv = tnf.oracle_system(X, conn, params, fixed, 1.0)
    _check_modes(lam, vecs, rp, ci, kv, mv)


@pytest.mark.parametrize("spectrum", ["lower", "upper"])
def test_clamped_block_against_dense(ctx, spectrum):
    X, conn, _ = problems.box_mesh(3, 4, 6, L=(0.6, 0.8, 2.0), jitter=0.15)
    rng = np.random.default_rng(9)
    params = np.stack([rng.uniform(2.0, 4.0, len(conn)), rng.uniform(0.8, 1.2, len(conn)),
                       np.zeros(len(conn)), np.zeros(len(conn))], axis=1)
    rho = rng.uniform(0.5, 1.5, len(conn))
    z0 = np.nonzero(X[:, 2] == 0)[0]
    fixed = np.sort(np.concatenate([3 * z0, 3 * z0 + 1, 3 * z0 + 2])).astype(np.int32)
    m = _device(X, conn, params, ctx)
    fm = neon_amd.natural_frequency_matrix(m, {"type": "arpack", "eigenvalues": 6, "spectrum": spectrum},
                                           [{"dofs": fixed}], rho)
    # the top of the spectrum is dense: a basis that nearly spans the space
    lam, vecs = fm.solve(max_basis=0 if spectrum == "lower" else 300)
    print(spectrum, "solves", fm.solves, lam, flush=True)
    rp, ci, kv, mv = tnf.oracle_system(X, conn, params, fixed, rho)
    ref = scipy.linalg.eigh(tnf._dense(rp, ci, kv), tnf._dense(rp, ci, mv), eigvals_only=True)
    want = ref[:6] if spectrum == "lower" else ref[-6:]
    np.testing.assert_allclose(lam, want, rtol=1e-8)
    _check_modes(lam, vecs, rp, ci, kv, mv)


def test_repeated_eigenvalues_of_a_square_cantilever(ctx):
    """A square-section cantilever: bending in x and y (and every mode pair related by the
    x <-> y symmetry) has double eigenvalues; the block solver returns both copies."""
    X, conn, _ = problems.box_mesh(2, 2, 8, L=(0.5, 0.5, 4.0))
    params = np.tile([3.0, 1.0, 0, 0], (len(conn), 1))
    z0 = np.nonzero(X[:, 2] == 0)[0]
    fixed = np.sort(np.concatenate([3 * z0, 3 * z0 + 1, 3 * z0 + 2])).astype(np.int32)
    m = _device(X, conn, params, ctx)
    fm = neon_amd.natural_frequency_matrix(m, {"type": "arpack", "eigenvalues": 8}, [{"dofs": fixed}], 2.0)
    lam, vecs = fm.solve()
    rp, ci, kv, mv = tnf.oracle_system(X, conn, params, fixed, 2.0)
    ref = scipy.linalg.eigh(tnf._dense(rp, ci, kv), tnf._dense(rp, ci, mv), eigvals_only=True)[:8]
    assert abs(ref[1] - ref[0]) < 1e-10 * ref[0]  # the first bending pair is double
    np.testing.assert_allclose(lam, ref, rtol=1e-8)
    _check_modes(lam, vecs, rp, ci, kv, mv)


def test_neohooke_at_rest_and_ingest(ctx, tmp_path):
    """Any model at the undeformed state: neo-Hooke's tangent at F = I is the small-strain one, so
    its frequencies equal linear elasticity's; and a natural-frequency step read by ingest."""
    mat = {"elastic_modulus": 10.0, "poissons_ratio": 0.3, "density": 2.0}
    p = ingest.read(tnf._frequency_json(tmp_path, mat, {"type": "arpack", "eigenvalues": 5}))
    lam = {}
    for model in (orc.MODEL_LINEAR_ELASTIC, orc.MODEL_NEOHOOKE):
        m = _device(p.X, p.conn, p.params, ctx, model=model)
        fm = neon_amd.natural_frequency_matrix(m, p.eigen_solver, p.boundaries, p.density)
        lam[model] = fm.solve()[0]
    np.testing.assert_allclose(lam[orc.MODEL_NEOHOOKE], lam[orc.MODEL_LINEAR_ELASTIC], rtol=1e-9)
    rp, ci, kv, mv = tnf.oracle_system(p.X, p.conn, p.params, p.fixed_dofs(), p.density)
    ref = scipy.linalg.eigh(tnf._dense(rp, ci, kv), tnf._dense(rp, ci, mv), eigvals_only=True)[:5]
    np.testing.assert_allclose(lam[orc.MODEL_LINEAR_ELASTIC], ref, rtol=1e-8)
    assert "radians / time" in fm.eigenvalue_table()


def test_cantilever_at_scale(ctx):
    """A 32 x 32 x 128-cell cantilever (421,443 dofs): the first bending pair is double and the
    modes satisfy K x = lambda M x; the time and the inner solves are printed."""
    X, conn, _ = problems.box_mesh(32, 32, 128, L=(1.0, 1.0, 4.0))
    params = np.tile([3.0, 1.0, 0, 0], (len(conn), 1))
    z0 = np.nonzero(X[:, 2] == 0)[0]
    fixed = np.sort(np.concatenate([3 * z0, 3 * z0 + 1, 3 * z0 + 2])).astype(np.int32)
    m = _device(X, conn, params, ctx)
    fm = neon_amd.natural_frequency_matrix(m, {"type": "arpack", "eigenvalues": 6}, [{"dofs": fixed}], 1.0)
    t0 = time.perf_counter()
    lam, vecs = fm.solve()
    print(f"cantilever 32x32x128: {time.perf_counter() - t0:.2f} s, {fm.solves} inner solves, lambda {lam}",
          flush=True)
    assert abs(lam[1] - lam[0]) < 1e-6 * lam[0]
    # Euler-Bernoulli first bending frequency of a cantilever as a sanity bound (shear and the
    # Poisson effect make the 3-D value a few per cent lower)
    E = 9 * 3.0 * 1.0 / (3 * 3.0 + 1.0)
    w_eb = 1.875104 ** 2 * np.sqrt(E * (1.0 / 12.0) / (1.0 * 1.0)) / 4.0 ** 2
    assert 0.85 < np.sqrt(lam[0]) / w_eb < 1.02
