"""Natural frequency analysis (natural_frequency_matrix, src/assembler/mechanics/
natural_frequency_matrix.hpp) in the oracle and the host API, on the CPU.

What pins it:
* the consistent mass (submesh::consistent_mass, submesh.cpp:162-183): on a box element
  M_ab = rho V / 64 prod_i (1 + xi_a,i xi_b,i / 3) per dof direction (the 2-point rule integrates
  the trilinear products exactly), and the assembled mass of a jittered unit cube sums to rho per
  direction;
* the generalised eigenproblem, solved densely (scipy.linalg.eigh) on the oracle's K and M after
  neon's homogeneous Dirichlet treatment: a bar with only its axial dofs free and its base fixed
  is exactly the 1-D linear-element chain with consistent mass, whose eigenvalues are
  lambda_k = 6 M_c / (rho h^2) (1 - cos t_k) / (2 + cos t_k), t_k = (2k - 1) pi / (2N), with the
  constrained modulus M_c = K + 4G/3;
* the eigen solver factory and the ingest of a natural-frequency step (neon's messages).
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg

import neon_amd
from neon_amd import ingest, problems
from oracle import oracle as orc

import neon_files


def _dense(rp, ci, val):
    n = len(rp) - 1
    A = np.zeros((n, n))
    for r in range(n):
        A[r, ci[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    return A


def bar(N=20, s=0.1, E=1.0, nu=0.3):
    """A 1 x 1 x N-cell bar of cross-section s x s and cell length 1: its params (K, G) and the
    fixed dofs that leave only the axial motion of the nodes above the base."""
    X, conn, _ = problems.box_mesh(1, 1, N, L=(s, s, float(N)))
    K, G = E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu))
    fixed = np.concatenate([3 * np.arange(len(X)), 3 * np.arange(len(X)) + 1,
                            3 * np.nonzero(X[:, 2] == 0)[0] + 2]).astype(np.int32)
    return X, conn, np.tile([K, G, 0.0, 0.0], (len(conn), 1)), np.unique(fixed)


def bar_eigenvalues(N, K, G, rho=1.0, h=1.0):
    t = (2 * np.arange(1, N + 1) - 1) * np.pi / (2 * N)
    return np.sort(6 * (K + 4 * G / 3) / (rho * h * h) * (1 - np.cos(t)) / (2 + np.cos(t)))


def oracle_system(X, conn, params, fixed, density, model=orc.MODEL_LINEAR_ELASTIC):
    m = orc.Model(X, conn, params, model=model)
    assert m.update(np.zeros(m.n_dofs)) == orc.OK
    rp, ci = m.pattern()
    kv, _ = m.assemble(rp, ci)
    mv = m.assemble_mass(rp, ci, density)
    orc.enforce_dirichlet(rp, ci, kv, np.zeros(m.n_dofs), fixed)
    orc.enforce_dirichlet(rp, ci, mv, np.zeros(m.n_dofs), fixed)
    return rp, ci, kv, mv


def test_element_mass_closed_form():
    X, conn, _ = problems.box_mesh(1, 1, 1, L=(0.5, 2.0, 3.0))
    m = orc.Model(X, conn, [[1.0, 1.0, 0, 0]], model=orc.MODEL_LINEAR_ELASTIC)
    me = m.element_mass(0, 7.9)
    xi = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                   [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)
    # the box mesh's node order within the element: map through the coordinates
    loc = (X[conn[0]] - X[conn[0]].min(axis=0)) / np.ptp(X[conn[0]], axis=0) * 2 - 1
    np.testing.assert_allclose(loc, xi, atol=1e-15)
    exact = 7.9 * 3.0 / 64 * np.prod(1 + xi[:, None, :] * xi[None, :, :] / 3, axis=2)
    for i in range(3):
        np.testing.assert_allclose(me[i::3, i::3], exact, rtol=1e-14)
        for j in range(3):
            if i != j:
                assert not me[i::3, j::3].any()


def test_assembled_mass_sums_to_the_mass():
    X, conn, _ = neon_files.cube_mesh(3)
    rng = np.random.default_rng(5)
    inner = np.all((X > 0) & (X < 1), axis=1)
    X[inner] += rng.uniform(-0.04, 0.04, (inner.sum(), 3))
    m = orc.Model(X, conn, np.tile([1.0, 1.0, 0, 0], (len(conn), 1)), model=orc.MODEL_LINEAR_ELASTIC)
    rp, ci = m.pattern()
    rho = rng.uniform(1.0, 2.0, len(conn))
    M = _dense(rp, ci, m.assemble_mass(rp, ci, rho))
    np.testing.assert_allclose(M, M.T, atol=1e-16)
    vols = np.array([m.element_mass(e, 1.0)[0::3, 0::3].sum() for e in range(len(conn))])
    assert vols.sum() == pytest.approx(1.0, rel=1e-13)  # the jitter keeps the cube
    for i in range(3):
        assert M[i::3, i::3].sum() == pytest.approx((rho * vols).sum(), rel=1e-13)
    assert np.linalg.eigvalsh(M).min() > 0


def test_bar_eigenvalues_closed_form():
    N = 20
    X, conn, params, fixed = bar(N)
    rp, ci, kv, mv = oracle_system(X, conn, params, fixed, 1.0)
    lam = scipy.linalg.eigh(_dense(rp, ci, kv), _dense(rp, ci, mv), eigvals_only=True)
    K, G = params[0, :2]
    np.testing.assert_allclose(lam[:10], bar_eigenvalues(N, K, G)[:10], rtol=1e-10)


def test_make_eigen_solver_messages():
    assert neon_amd.make_eigen_solver({"type": "arpack"}) == {"type": "arpack", "values": 10,
                                                             "spectrum": neon_amd.SPECTRUM_LOWER}
    assert neon_amd.make_eigen_solver({"type": "arpack", "eigenvalues": 7, "spectrum": "upper"})["values"] == 7
    for data, msg in (({}, "type was not provided"), ({"type": "arpack", "eigenvalues": 0}, "must be positive"),
                      ({"type": "arpack", "spectrum": "middle"}, '"lower" and "upper"'),
                      ({"type": "lanczos"}, "OpenCL"), ({"type": "power_iteration"}, "OpenCL"),
                      ({"type": "slepc"}, "not recognised")):
        with pytest.raises(ValueError, match=msg):
            neon_amd.make_eigen_solver(data)


def _frequency_json(tmp_path, material, eigen=None, solution="natural_frequency"):
    path = neon_files.write_cube_simulation(tmp_path, 2, material, {"name": "isotropic_linear_elasticity"},
                                            [{"name": "z0", "type": "displacement", "time": [0.0, 1.0],
                                              "x": [0.0, 0.0], "y": [0.0, 0.0], "z": [0.0, 0.0]}])
    with open(path) as fh:
        sim = json.load(fh)
    step = sim["steps"][0]
    step["solution"] = solution
    del step["linear_solver"]
    if eigen is not None:
        step["eigen_solver"] = eigen
    with open(path, "w") as fh:
        json.dump(sim, fh)
    return path


def test_ingest_natural_frequency_step(tmp_path):
    mat = {"elastic_modulus": 200.0e9, "poissons_ratio": 0.3, "density": 7900.0}
    p = ingest.read(_frequency_json(tmp_path / "a", mat, {"type": "arpack", "eigenvalues": 7, "spectrum": "lower"}))
    assert p.solution == "natural_frequency" and p.eigen_solver["eigenvalues"] == 7
    np.testing.assert_array_equal(p.density, np.full(8, 7900.0))
    assert len(p.fixed_dofs()) == 27
    with pytest.raises(ValueError, match="density was requested"):
        ingest.read(_frequency_json(tmp_path / "b", {"elastic_modulus": 1.0, "poissons_ratio": 0.3},
                                    {"type": "arpack"}))
    with pytest.raises(ValueError, match="eigen_solver"):
        ingest.read(_frequency_json(tmp_path / "c", mat))
    with pytest.raises(ValueError, match='"solution" is not valid'):
        ingest.read(_frequency_json(tmp_path / "d", mat, {"type": "arpack"}, solution="explicit_dynamics"))
    assert os.path.exists(tmp_path / "a" / "cube.mesh")
