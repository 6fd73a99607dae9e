"""CPU checks of the C ABI library: it loads, exports every function that
include/neon_hip.h declares, and its host-only parts (mesh generator, random streams,
argument checks) behave.  No compute call needs a GPU here."""
import pathlib
import re

import numpy as np
import pytest

from neon_amd import _capi, problems

ROOT = pathlib.Path(__file__).resolve().parent.parent


def declared_functions(header="neon_hip.h"):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(neon_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _capi.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    # and the Python binding declares a signature for each
    assert set(names) <= set(_capi.SIGNATURES)


def test_probes_live_in_the_measurement_library_only():
    # the timing probes are exported by libneon_hip_measure.so, which loads against the
    # product library, and not by the product library itself
    names = declared_functions("neon_hip_measure.h")
    assert len(names) == 4
    measure = _capi.measure_lib()
    assert all(hasattr(measure, n) for n in names)
    assert set(names) == set(_capi.MEASURE_SIGNATURES)
    product = _capi.lib()
    assert not any(hasattr(product, n) for n in names)


def test_no_device_is_a_loud_error_here():
    if _capi.device_count() > 0:
        pytest.skip("a GPU is present")
    import neon_amd

    with pytest.raises(neon_amd.NeonError) as e:
        neon_amd.Context(0)
    assert e.value.code == _capi.NEON_NO_DEVICE


def test_box_mesh_structure_and_orientation():
    X, conn, ijk = problems.box_mesh(3, 4, 5, L=(3.0, 4.0, 5.0), jitter=0.1)
    assert X.shape == (4 * 5 * 6, 3) and conn.shape == (60, 8)
    n = 1 + 4 * (1 + 5 * 1)  # node (1,1,1)
    assert tuple(ijk[n]) == (1, 1, 1)
    # every element has positive Jacobian at its centre and nodes in hex8 order
    for e in conn:
        x = X[e]
        J = np.stack([(x[1] - x[0]), (x[3] - x[0]), (x[4] - x[0])])
        assert np.linalg.det(J) > 0
    # jitter stays within 10% of h and never moves a node off its boundary face
    base = ijk * np.array([1.0, 1.0, 1.0])
    assert np.abs(X - base).max() <= 0.1 + 1e-15
    for c, nmax in enumerate((3, 4, 5)):
        on = (ijk[:, c] == 0) | (ijk[:, c] == nmax)
        np.testing.assert_array_equal(X[on, c], base[on, c])


def test_box_mesh_is_deterministic_and_seeded():
    a = problems.box_mesh(4, 4, 4, jitter=0.1, seed=1)[0]
    b = problems.box_mesh(4, 4, 4, jitter=0.1, seed=1)[0]
    c = problems.box_mesh(4, 4, 4, jitter=0.1, seed=2)[0]
    np.testing.assert_array_equal(a, b)
    assert np.abs(a - c).max() > 0


def test_box_mesh_hole():
    nx, ny, nz = 16, 8, 2
    h = 1.0 / ny
    X, conn, ijk = problems.box_mesh(nx, ny, nz, L=(2.0, 1.0, nz * h), hole_radius=2 * h)
    centres = X[conn].mean(axis=1)
    r = np.hypot(centres[:, 0] - 1.0, centres[:, 1] - 0.5)
    assert (r > 2 * h).all()
    assert conn.shape[0] < nx * ny * nz
    assert len(np.unique(conn)) == X.shape[0]  # unused nodes dropped


def test_random_streams():
    u = problems.uniform(problems.SEED_MU, 200000, 0.5, 1.5)
    assert u.min() >= 0.5 and u.max() < 1.5
    assert u.mean() == pytest.approx(1.0, abs=0.005)
    z = problems.normal(problems.SEED_LOGMU, 200000)
    assert z.mean() == pytest.approx(0.0, abs=0.01) and z.std() == pytest.approx(1.0, abs=0.01)
    np.testing.assert_array_equal(problems.uniform(7, 10, 0, 1), problems.uniform(7, 10, 0, 1)[:10])


def test_problem_definitions():
    p = problems.cube_clamped(n=10)
    assert (p.n_elem, p.n_nodes, p.n_dofs) == (1000, 1331, 3993)
    p = problems.block_rollers(8)
    mu = p.params[:, 1]
    assert mu.min() >= 0.5 and mu.max() <= 1.5
    np.testing.assert_allclose(p.params[:, 0], 10 * mu)
    p = problems.notched_plate_j2(16, 8, 2)
    assert p.params[:, 2].min() >= 225 and p.params[:, 2].max() <= 275


def test_make_linear_solver_selection():
    import neon_amd

    with pytest.raises(ValueError):
        neon_amd.make_linear_solver({}, True)
    with pytest.raises(ValueError):
        neon_amd.make_linear_solver({"type": "iterative", "device": "gpu"}, True)
    with pytest.raises(ValueError):
        neon_amd.make_linear_solver({"type": "iterative", "device": "gpu", "backend": {"type": "cuda"}}, True)


def test_model_traits_follow_the_reference_constructors():
    # constitutive_model::is_finite_deformation / is_symmetric / names of each model
    # (compressible_neohooke.hpp:39, small_strain_J2_plasticity.hpp:32 and .cpp:19-23,
    # isotropic_linear_elasticity.cpp:17-20, gaussian_ageing_affine_microsphere.cpp:30-34)
    import neon_amd

    assert neon_amd.model_traits(neon_amd.MODEL_NEOHOOKE) == (True, True)
    assert neon_amd.model_traits(neon_amd.MODEL_J2) == (False, True)
    assert neon_amd.model_traits(neon_amd.MODEL_LINEAR_ELASTIC) == (False, True)
    for m in (neon_amd.MODEL_MICROSPHERE_GAUSSIAN, neon_amd.MODEL_MICROSPHERE_LANGEVIN,
              neon_amd.MODEL_MICROSPHERE_NONAFFINE, neon_amd.MODEL_MICROSPHERE_AGEING,
              neon_amd.MODEL_MICROSPHERE_GAUSSIAN | neon_amd.SPHERE_BO33):
        assert neon_amd.model_traits(m) == (True, True)
    names = _capi.lib().neon_model_variable_names(neon_amd.MODEL_J2).decode().split(";")
    assert names == sorted(["linearised_strain", "von_mises_stress", "linearised_plastic_strain",
                            "effective_plastic_strain"])
    assert _capi.lib().neon_model_variable_names(neon_amd.MODEL_NEOHOOKE).decode() == ""
    assert _capi.lib().neon_model_variable_names(99) is None
    with pytest.raises(neon_amd.NeonError):
        neon_amd.model_traits(99)
