"""The LATIN iterations (latin_matrix, src/assembler/mechanics/latin_matrix.hpp:150-244, with
latin_submesh::incremental_latin_internal_force, src/mesh/mechanics/solid/latin_submesh.hpp:62-111)
in the oracle, the host API and ingest, on the CPU.

What pins the oracle's LATIN force f(s) = sum_e int B^T (sigma_old + s D (eps - eps_old) - sigma):
* at the committed state (u = u_old, sigma = sigma_old) it is exactly zero for every s;
* for linear elasticity sigma = D eps, so f(1) = 0 and f(s) = (1 - s) f(0) with f(0) != 0: a
  wrong Voigt factor, index or sign in applying D, or old and current swapped, breaks this;
* with s = 0 the right-hand side is int B^T (sigma_old - sigma), which equals the Newton
  residual f_ext - f_int up to the committed state's own residual, so on the reference's J2
  cube (zero external force) the LATIN iterations reach static_matrix's solution.
The reference's own LATIN test (test/matrix_system.cpp:67-91) only runs the solver."""
import copy

import numpy as np
import pytest

import neon_amd
from neon_amd import ingest, problems
from oracle import oracle as orc

import neon_files
import test_ingest as ti


def _block(model, params):
    X, conn, _ = problems.box_mesh(3, 3, 3, L=(1.0, 1.0, 1.0), jitter=0.15)
    return orc.Model(X, conn, np.tile(params, (len(conn), 1)), model=model, tangent=orc.TANGENT_REFERENCE), X


def _u(X, scale, seed):
    rng = np.random.default_rng(seed)
    return scale * (X @ rng.uniform(-1, 1, (3, 3)).T + 0.3 * np.sin(3 * X)).reshape(-1)


@pytest.mark.parametrize("model,params", [(orc.MODEL_J2, [3.0e3, 1.0e3, 5.0, 50.0]),
                                          (orc.MODEL_NEOHOOKE, [3.0, 1.0, 0.0, 0.0]),
                                          (orc.MODEL_MICROSPHERE_LANGEVIN, [3.0, 1.0, 20.0, 0.0])])
def test_latin_force_vanishes_at_the_committed_state(model, params):
    m, X = _block(model, params)
    assert m.update(_u(X, 0.02, 1)) == orc.OK
    m.save(True)
    for s in (0.0, 0.7, 1.0, 2.5):
        assert not m.latin_force(s).any()


def test_latin_force_of_linear_elasticity():
    m, X = _block(orc.MODEL_LINEAR_ELASTIC, [3.0, 1.0, 0.0, 0.0])
    assert m.update(_u(X, 1e-3, 2)) == orc.OK
    m.save(True)
    assert m.update(_u(X, 2e-3, 3)) == orc.OK
    f0, f1, fh = m.latin_force(0.0), m.latin_force(1.0), m.latin_force(0.25)
    scale = np.abs(f0).max()
    assert scale > 1e-4
    assert np.abs(f1).max() <= 1e-12 * scale
    np.testing.assert_allclose(fh, 0.75 * f0, rtol=0, atol=1e-12 * scale)


def test_latin_with_zero_direction_reaches_the_newton_solution():
    p = ingest.read(ti.J2_CUBE)
    m0 = orc.Model(p.X, p.conn, p.params, model=p.model, tangent=orc.TANGENT_REFERENCE)
    code0, u0, recs0 = m0.newton(orc.newton_options(p.simulation), p.boundaries, loads=p.loads)
    m = orc.Model(p.X, p.conn, p.params, model=p.model, tangent=orc.TANGENT_REFERENCE)
    code, u, recs = m.newton(orc.newton_options(p.simulation, latin=0.0), p.boundaries, loads=p.loads)
    assert code0 == orc.OK and code == orc.OK
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-10 * np.abs(u0).max())
    np.testing.assert_allclose(m.qp_state()["ebar"], m0.qp_state()["ebar"], rtol=1e-8)
    # the first iteration is Newton's: the same first residual
    assert recs[0]["force_norm"] == pytest.approx(recs0[0]["force_norm"], rel=1e-12)


def test_latin_elastic_direction_stalls():
    """With s = 1 an elastic point's LATIN stress equals its constitutive stress, so the plastic
    J2 cube's iterations stop moving and the load step is cut back (neon behaves the same: its
    only LATIN test runs the solver without checking a result)."""
    p = ingest.read(ti.J2_CUBE)
    m = orc.Model(p.X, p.conn, p.params, model=p.model, tangent=orc.TANGENT_REFERENCE)
    code, u, recs = m.newton(orc.newton_options(p.simulation, latin=1.0), p.boundaries, loads=p.loads)
    assert code != orc.OK
    assert recs[1]["step_time"] == 1.0 and recs[10]["step_time"] == 0.5  # ten iterations, then a cut-back


def _latin_json(tmp_path, nonlinear):
    path = neon_files.write_cube_simulation(tmp_path, 2, {"elastic_modulus": 10.0, "poissons_ratio": 0.3},
                                            {"name": "neohooke"}, neon_files.rollers(
                                                [{"name": "z1", "type": "displacement", "time": [0.0, 1.0],
                                                  "z": [0.0, 0.1]}]), nonlinear=nonlinear)
    import json
    with open(path) as fh:
        sim = json.load(fh)
    sim["steps"][0]["solution"] = "latin"
    with open(path, "w") as fh:
        json.dump(sim, fh)
    return path


def test_ingest_latin_step(tmp_path):
    base = {"residual_tolerance": 1e-8, "displacement_tolerance": 1e-8}
    ok = dict(base, latin_scheme="incremental", latin_local_stage="implicit", latin_search_direction=0.0)
    p = ingest.read(_latin_json(tmp_path / "a", ok))
    assert p.solution == "latin" and p.simulation["nonlinear_options"]["latin_search_direction"] == 0.0
    with pytest.raises(ValueError, match="Only incremental and implicit"):
        ingest.read(_latin_json(tmp_path / "b", dict(ok, latin_local_stage="explicit")))
    with pytest.raises(ValueError, match="latin_search_direction not specified"):
        ingest.read(_latin_json(tmp_path / "c", dict(base, latin_scheme="incremental", latin_local_stage="implicit")))
    sim = copy.deepcopy(p.simulation)
    del sim["nonlinear_options"]["latin_search_direction"]
    class _host_only_mesh:  # latin_matrix checks its options after static_matrix's constructor
        def is_symmetric(self):
            return True

    with pytest.raises(ValueError, match="latin_search_direction not specified"):
        neon_amd.latin_matrix(_host_only_mesh(), sim, p.boundaries)
