"""Problem ingest (neon's simulation JSON + .mesh) and nonfollower loads, on the CPU.

* The reference's own example inputs (tests/golden/neon_examples: the neo-Hooke cube of
  examples/solid_mechanics/hyperelastic/neohooke/cube and the J2 linear-hex cube of
  examples/solid_mechanics/hypoelastic_plastic/cube/linear-hex, input files copied as data)
  are read and solved by the oracle, and the solutions checked against what fixes them:
  verify.py's checks for the neo-Hooke cube (maximum u_z 0.5, u_z linear along the z axis
  with slope 0.5, verify.py:85-93) plus the lateral stretch of uniaxial neo-Hooke stress
  (lambda ln J + mu (l^2 - 1) = 0, J = l^2 l_z), and the 1-D J2 return map for the J2 cube
  (sigma = (sigma_y + H eps) E / (E + H), eps_p = eps - sigma / E, lateral strain
  -nu sigma / E - eps_p / 2).
* The library's unit-load vectors (C ABI, host only) against the oracle's own numpy
  restatement, and both against closed forms (face areas, pressure resultants, volumes).
* Traction, pressure, body force and nodal force through ingest and the oracle's Newton:
  homogeneous uniaxial neo-Hooke under a traction or pressure (P_zz = t with the lateral
  equation above) and the reactions balancing the applied body and nodal forces.
"""
import os

import numpy as np
import pytest
from scipy.optimize import brentq, fsolve

import neon_amd
from neon_amd import _capi, ingest, problems
from oracle import oracle as orc

import neon_files

HERE = os.path.dirname(os.path.abspath(__file__))
NEOHOOKE_CUBE = os.path.join(HERE, "golden", "neon_examples", "neohooke_cube", "cube.json")
J2_CUBE = os.path.join(HERE, "golden", "neon_examples", "j2_cube", "cube.json")
DAMAGE_CUBE = os.path.join(HERE, "golden", "neon_examples", "viscoplastic_damage_cube", "cube.json")


def oracle_solve(p, threads=0):
    m = orc.Model(p.X, p.conn, p.params, model=p.model, tangent=orc.TANGENT_REFERENCE)
    damage = getattr(p, "damage", None)
    if damage is not None:
        m.set_damage(damage)
    # the unsymmetric tangent of J2 with damage is solved with BiCGSTAB, as make_linear_solver
    # picks for a model whose is_symmetric() is false
    solver = 1 if p.model == orc.MODEL_J2_DAMAGE else 0
    code, u, recs = m.newton(orc.newton_options(p.simulation, threads=threads, solver=solver), p.boundaries,
                             loads=p.loads)
    return code, u, recs, m


def uniaxial_neohooke_lateral(K, G, lz):
    lam = K - 2.0 * G / 3.0
    return brentq(lambda l: lam * np.log(l * l * lz) + G * (l * l - 1.0), 0.3, 1.5, xtol=1e-15)


def uniaxial_neohooke_traction(K, G, t):
    """(l_z, l_lat) with first Piola stress P_zz = t and no lateral stress."""
    lam = K - 2.0 * G / 3.0

    def eqs(v):
        lz, ll = v
        lnJ = np.log(ll * ll * lz)
        return [lam * lnJ + G * (ll * ll - 1.0), (lam * lnJ + G * (lz * lz - 1.0)) / lz - t]

    return fsolve(eqs, [1.0, 1.0], xtol=1e-13)


# ------------------------------------------------------------------ the reference's examples
def test_ingest_neohooke_cube_example():
    p = ingest.read(NEOHOOKE_CUBE)
    assert (p.n_nodes, p.n_elem, p.model) == (216, 125, _capi.MODEL_NEOHOOKE)
    E, nu = 10.0e6, 0.49
    assert p.params[0, 0] == pytest.approx(E / (3 * (1 - 2 * nu)), rel=1e-15)
    assert p.params[0, 1] == pytest.approx(E / (2 * (1 + nu)), rel=1e-15)
    assert [(b["name"], len(b["dofs"])) for b in p.boundaries] == [("Xsym", 36), ("Ysym", 36), ("Zsym", 36),
                                                                 ("ZLoad", 36)]
    # Xsym / Ysym are the x = 1 / y = 1 faces in this mesh; ZLoad the top, displaced to 0.5
    assert np.all(p.X[np.asarray(p.boundaries[0]["dofs"]) // 3, 0] == 1.0)
    assert np.all(np.asarray(p.boundaries[0]["dofs"]) % 3 == 0)
    assert list(p.boundaries[3]["values"]) == [0.0, 0.5]
    assert p.simulation["linear_solver"] == {"tolerance": 1e-8, "maximum_iterations": 2000}
    assert p.loads == []


def test_oracle_neohooke_cube_example_closed_form():
    p = ingest.read(NEOHOOKE_CUBE)
    code, u, recs, _ = oracle_solve(p)
    assert code == orc.OK
    U = u.reshape(-1, 3)
    # verify.py: maximum displacement 0.5 and u_z = 0.5 z along the z axis
    assert U.max() == pytest.approx(0.5, rel=1e-9)
    axis = np.nonzero((p.X[:, 0] == 0) & (p.X[:, 1] == 0))[0]
    slope, icpt = np.polyfit(p.X[axis, 2], U[axis, 2], 1)
    assert slope == pytest.approx(0.5, rel=1e-7) and abs(icpt) < 1e-7
    # uniaxial stress: lateral stretch towards the x = 1, y = 1 symmetry faces; the example's
    # own Newton tolerance is 1e-3 (relative), which leaves ~2e-7 here
    ll = uniaxial_neohooke_lateral(p.params[0, 0], p.params[0, 1], 1.5)
    np.testing.assert_allclose(U[:, 0], (ll - 1.0) * (p.X[:, 0] - 1.0), atol=1e-5)
    np.testing.assert_allclose(U[:, 1], (ll - 1.0) * (p.X[:, 1] - 1.0), atol=1e-5)
    np.testing.assert_allclose(U[:, 2], 0.5 * p.X[:, 2], atol=1e-8)


def test_oracle_j2_cube_example_closed_form():
    p = ingest.read(J2_CUBE)
    assert p.model == _capi.MODEL_J2
    # PaStiX has no device counterpart: CG to a relative 1e-12
    assert p.simulation["linear_solver"]["tolerance"] == 1e-12
    code, u, recs, m = oracle_solve(p)
    assert code == orc.OK
    K, G, sy, H = p.params[0]
    E = 9 * K * G / (3 * K + G)
    nu = (3 * K - 2 * G) / (2 * (3 * K + G))
    eps = 1.0e-3
    sig = (sy + H * eps) * E / (E + H)
    ep = eps - sig / E
    elat = -nu * sig / E - 0.5 * ep
    U = u.reshape(-1, 3)
    np.testing.assert_allclose(U[:, 2], eps * p.X[:, 2], rtol=0, atol=1e-13)
    np.testing.assert_allclose(U[:, 0], elat * (p.X[:, 0] - 1.0), rtol=0, atol=1e-13)
    np.testing.assert_allclose(U[:, 1], elat * (p.X[:, 1] - 1.0), rtol=0, atol=1e-13)
    np.testing.assert_allclose(m.qp_state()["ebar"], ep, rtol=1e-9)


def test_ingest_rejects_what_the_device_path_does_not_run(tmp_path):
    mat = {"elastic_modulus": 10.0, "poissons_ratio": 0.3}
    nh = {"name": "neohooke"}
    bad = [
        (mat, {"name": "microsphere"}, [], 'Missing "type"'),
        (mat, {"name": "microsphere", "type": "affine"}, [], 'Missing "quadrature"'),
        (mat, {"name": "microsphere", "type": "affine", "quadrature": "BO61"}, [], '"quadrature" field must be'),
        (mat, {"name": "microsphere", "type": "affine", "quadrature": "BO21"}, [], 'Missing "statistics"'),
        (mat, {"name": "microsphere", "type": "affine", "quadrature": "BO21", "statistics": "x"}, [],
         '"statistics" for the microsphere'),
        (mat, {"name": "microsphere", "type": "affine", "quadrature": "BO21", "statistics": "gaussian",
               "ageing": "other"}, [], "BAND"),
        (mat, {"name": "microsphere", "type": "affine", "quadrature": "FM900", "statistics": "gaussian"}, [],
         "FM900"),
        (dict(mat, segments_per_chain=20, recombination_probability=1e-5, cure_time=80),
         {"name": "microsphere", "type": "affine", "quadrature": "BO21", "statistics": "gaussian", "ageing": "BAND"},
         [], '"scission_probability" is not specified'),
        (dict(mat, segments_per_chain=20, scission_probability=1e-5, recombination_probability=1e-5, cure_time=10),
         {"name": "microsphere", "type": "affine", "quadrature": "BO21", "statistics": "gaussian", "ageing": "BAND"},
         [], "less than 10"),
        (dict(mat, segments_per_chain=20, scission_probability=-1e-5, recombination_probability=1e-5, cure_time=50),
         {"name": "microsphere", "type": "affine", "quadrature": "BO21", "statistics": "gaussian", "ageing": "BAND"},
         [], "must be positive"),
        (mat, {"name": "microsphere", "type": "cubic", "quadrature": "BO21"}, [], '"affine" or "nonaffine"'),
        (mat, {"name": "microsphere", "type": "affine", "quadrature": "BO21", "statistics": "gaussian"}, [],
         "segments_per_chain"),
        (dict(mat, segments_per_chain=20), {"name": "microsphere", "type": "nonaffine", "quadrature": "BO21"}, [],
         "nonaffine_stretch_parameter"),
        (mat, {"name": "mooney_rivlin"}, [], "is not recognised"),
        (mat, {"name": "J2_plasticity"}, [], "finite_strain"),
        (mat, {"name": "J2_plasticity", "finite_strain": True}, [], "finite-strain J2"),
        (mat, {"name": "J2_plasticity", "finite_strain": True, "damage": "isotropic_chaboche"}, [],
         "not yet implemented"),
        (dict(mat, yield_stress=85.0), {"name": "J2_plasticity", "finite_strain": False, "damage": "isotropic_chaboche"},
         [], '"softening_multiplier" needs to be specified'),
        ({"elastic_modulus": 10.0, "poissons_ratio": 0.5}, nh, [], "poissons_ratio"),
        ({"elastic_modulus": 10.0}, nh, [], "bulk_modulus"),
        (mat, nh, [{"name": "z1", "type": "heat_flux", "time": [0, 1], "z": [0, 1]}], "not recognised"),
        (mat, nh, [{"name": "z1", "type": "pressure", "time": [0, 1]}], '"value"'),
        (mat, nh, [{"name": "nowhere", "type": "displacement", "time": [0, 1], "z": [0, 1]}], "does not exist"),
        (mat, nh, [{"name": "z1", "type": "displacement", "z": [0, 1]}], "time"),
    ]
    for k, (material, constitutive, bcs, msg) in enumerate(bad):
        path = neon_files.write_cube_simulation(tmp_path / str(k), 2, material, constitutive, bcs)
        with pytest.raises(ValueError, match=msg):
            ingest.read(path)
    # a body force on quadrilateral faces, a traction on hex8 cells
    for k, bc in enumerate([{"name": "z1", "type": "body_force", "time": [0, 1], "z": [0, 1]},
                            {"name": "cube", "type": "traction", "time": [0, 1], "z": [0, 1]}]):
        path = neon_files.write_cube_simulation(tmp_path / f"t{k}", 2, mat, nh, [bc])
        with pytest.raises(ValueError, match="need"):
            ingest.read(path)


def test_ingest_constitutive_models(tmp_path):
    mat = {"elastic_modulus": 10.0, "poissons_ratio": 0.3, "segments_per_chain": 25,
           "nonaffine_stretch_parameter": 1.5}
    K, G = 10.0 / (3 * 0.4), 10.0 / 2.6
    cases = [
        ({"name": "isotropic_linear_elasticity"}, _capi.MODEL_LINEAR_ELASTIC, [K, G, 0, 0]),
        ({"name": "microsphere", "type": "affine", "statistics": "gaussian", "quadrature": "BO21"},
         _capi.MODEL_MICROSPHERE_GAUSSIAN, [K, G, 25, 0]),
        ({"name": "microsphere", "type": "affine", "statistics": "langevin", "quadrature": "BO33"},
         _capi.MODEL_MICROSPHERE_LANGEVIN | _capi.SPHERE_BO33, [K, G, 25, 0]),
        ({"name": "microsphere", "type": "nonaffine", "quadrature": "BO33"},
         _capi.MODEL_MICROSPHERE_NONAFFINE | _capi.SPHERE_BO33, [K, G, 25, 1.5]),
    ]
    for k, (constitutive, model, prm) in enumerate(cases):
        p = ingest.read(neon_files.write_cube_simulation(tmp_path / str(k), 2, mat, constitutive, []))
        assert p.model == model
        np.testing.assert_allclose(p.params, np.tile(prm, (8, 1)), rtol=1e-15)


# ------------------------------------------------------------------ unit-load vectors
def _jittered_cube(n, seed):
    X, conn, faces = neon_files.cube_mesh(n)
    rng = np.random.default_rng(seed)
    interior = np.all((X > 0) & (X < 1), axis=1)
    X = X.copy()
    X[interior] += rng.uniform(-0.15 / n, 0.15 / n, size=(interior.sum(), 3))
    return X, conn, faces


def test_load_vectors_library_against_oracle():
    X, conn, faces = _jittered_cube(4, 3)
    # a warped face set: lift the top face's interior nodes
    Xw = X.copy()
    top = np.unique(faces["z1"])
    inner = top[(X[top, 0] > 0) & (X[top, 0] < 1) & (X[top, 1] > 0) & (X[top, 1] < 1)]
    Xw[inner, 2] += np.random.default_rng(4).uniform(-0.05, 0.05, len(inner))
    for Xc in (X, Xw):
        for name in ("z1", "x0", "y1"):
            for comp in (-1, 0, 1, 2):
                for red in (False, True):
                    a = neon_amd.surface_load_vector(Xc, faces[name], comp, red)
                    b = orc.surface_load_vector(Xc, faces[name], comp, red)
                    np.testing.assert_allclose(a, b, rtol=0, atol=1e-15)
        for comp in (0, 1, 2):
            for red in (False, True):
                np.testing.assert_allclose(neon_amd.volume_load_vector(Xc, conn, comp, red),
                                           orc.volume_load_vector(Xc, conn, comp, red), rtol=0, atol=1e-15)


def test_load_vectors_closed_forms():
    X, conn, faces = neon_files.cube_mesh(3)
    # a 2 x 3 x 5 box: faces of area 6, 10, 15, volume 30
    Xs = X * np.array([2.0, 3.0, 5.0])
    for name, area, normal in (("z1", 6.0, [0, 0, 1]), ("z0", 6.0, [0, 0, -1]), ("x1", 15.0, [1, 0, 0]),
                               ("x0", 15.0, [-1, 0, 0]), ("y1", 10.0, [0, 1, 0]), ("y0", 10.0, [0, -1, 0])):
        t = orc.surface_load_vector(Xs, faces[name], 1).reshape(-1, 3)
        assert t[:, 1].sum() == pytest.approx(area, rel=1e-14)
        assert np.abs(t[:, [0, 2]]).max() == 0.0
        # a uniform face of rectangles: 1/4 of each adjacent rectangle per node
        cell = area / 9.0
        used = np.unique(faces[name])
        counts = np.bincount(faces[name].ravel(), minlength=len(X))[used]
        np.testing.assert_allclose(t[used, 1], counts * cell / 4.0, rtol=1e-14)
        # unit pressure pushes against the outward normal with resultant -area n
        pr = neon_amd.surface_load_vector(Xs, faces[name], -1).reshape(-1, 3)
        np.testing.assert_allclose(pr.sum(axis=0), -area * np.array(normal), atol=1e-13)
    b = neon_amd.volume_load_vector(Xs, conn, 2).reshape(-1, 3)
    assert b[:, 2].sum() == pytest.approx(30.0, rel=1e-14)
    # a jittered mesh keeps the total volume (the jitter moves interior nodes only)
    Xj, conn, _ = _jittered_cube(3, 1)
    assert neon_amd.volume_load_vector(Xj, conn, 0).sum() == pytest.approx(1.0, rel=1e-13)


def test_load_vector_arguments_are_checked():
    X, conn, faces = neon_files.cube_mesh(2)
    out = np.zeros(3 * len(X))
    L = _capi.lib()
    bad_faces = faces["z1"].copy()
    bad_faces[0, 0] = len(X)
    assert L.neon_surface_load_vector(len(X), _capi.f64(X), len(bad_faces), _capi.i32(bad_faces.astype(np.int32)),
                                      0, 0, _capi.f64(out)) == _capi.NEON_BAD_ARGUMENT
    c32 = conn.astype(np.int32)
    assert L.neon_volume_load_vector(len(X), _capi.f64(X), len(c32), _capi.i32(c32), 3, 0,
                                     _capi.f64(out)) == _capi.NEON_BAD_ARGUMENT
    assert not out.any()


# ------------------------------------------------------------------ loads through the oracle's Newton
NEOHOOKE = {"bulk_modulus": 5.0, "shear_modulus": 1.0}


def _traction_problem(tmp_path, kind, value, n=3):
    if kind == "traction":
        bc = {"name": "z1", "type": "traction", "time": [0.0, 1.0], "z": [0.0, value]}
    else:
        bc = {"name": "z1", "type": "pressure", "time": [0.0, 1.0], "value": [0.0, value]}
    path = neon_files.write_cube_simulation(tmp_path, n, NEOHOOKE, {"name": "neohooke"}, neon_files.rollers([bc]))
    return ingest.read(path)


@pytest.mark.parametrize("kind, value, t", [("traction", 0.4, 0.4), ("pressure", -0.4, 0.4), ("pressure", 0.3, -0.3)])
def test_oracle_uniaxial_load_closed_form(tmp_path, kind, value, t):
    # a pressure acts against the outward normal: a negative pressure on the top face pulls
    p = _traction_problem(tmp_path, kind, value)
    assert len(p.loads) == 1 and p.boundaries
    code, u, recs, _ = oracle_solve(p)
    assert code == orc.OK
    lz, ll = uniaxial_neohooke_traction(5.0, 1.0, t)
    U = u.reshape(-1, 3)
    np.testing.assert_allclose(U[:, 2], (lz - 1.0) * p.X[:, 2], rtol=0, atol=1e-9)
    np.testing.assert_allclose(U[:, 0], (ll - 1.0) * p.X[:, 0], rtol=0, atol=1e-9)
    # the load history times step the solution: 4 equal increments of 0.25
    assert sorted({round(r["step_time"], 12) for r in recs}) == [0.25, 0.5, 0.75, 1.0]


def test_oracle_body_and_nodal_force_equilibrium(tmp_path):
    # body force in -z and a nodal force on the top face's nodes in +z; the nodal force is
    # applied at its t = 1 value from the first step (neon's compute_external_force)
    bcs = neon_files.rollers([
        {"name": "cube", "type": "body_force", "time": [0.0, 1.0], "z": [0.0, -0.3]},
        {"name": "z1", "type": "nodal_force", "time": [0.0, 1.0], "z": [0.0, 0.01]},
    ])
    path = neon_files.write_cube_simulation(tmp_path, 3, NEOHOOKE, {"name": "neohooke"}, bcs)
    p = ingest.read(path)
    assert [ld.get("fixed_time", -1.0) for ld in p.loads] == [-1.0, 1.0]
    code, u, recs, m = oracle_solve(p)
    assert code == orc.OK
    rp, ci = m.pattern()
    assert m.update(u) == orc.OK
    _, f = m.assemble(rp, ci)
    fext = np.zeros(p.n_dofs)
    for ld in p.loads:
        fext[ld["dofs"]] += ld["values"] * ld["loads"][-1]
    free = np.ones(p.n_dofs, dtype=bool)
    free[p.fixed_dofs()] = False
    # converged: f_int = f_ext on the free dofs; the z reactions f_int - f_ext on z0 balance
    # the applied loads (f_int sums to zero over all dofs)
    np.testing.assert_allclose(f[free], fext[free], rtol=0, atol=1e-9)
    z0 = np.asarray(p.boundaries[2]["dofs"])
    assert (f[z0] - fext[z0]).sum() == pytest.approx(-fext[2::3].sum(), rel=1e-9)
    assert fext[2::3].sum() == pytest.approx(-0.3 + 16 * 0.01, rel=1e-12)


def _sinusoid_bc(**kw):
    b = {"name": "z1", "type": "displacement", "generate_type": "sinusoidal"}
    b.update(kw)
    return b


def test_generated_sinusoidal_history(tmp_path):
    """boundary::generate_sinusoidal (boundary.cpp:87-143) with the first increment as its step:
    two blocks, the second shifted to start where the first ends, each value
    a_i sin(2 pi t / T_i + phase_i) at its absolute time; neon's messages for missing fields."""
    time = {"period": 2.0, "increments": {"initial": 0.25, "minimum": 0.01, "maximum": 0.25, "adaptive": False}}
    bc = _sinusoid_bc(z=[0.01, 0.02], period=[1.0, 0.5], number_of_cycles=[1.0, 2.0], phase=[0.0, np.pi / 2])
    path = neon_files.write_cube_simulation(tmp_path / "a", 2, NEOHOOKE, {"name": "neohooke"},
                                            neon_files.rollers([bc]), time=time)
    p = ingest.read(path)
    (b,) = [b for b in p.boundaries if b["name"] == "z1"]
    t = np.array([0.0, 0.25, 0.5, 0.75, 1.0, 1.25, 1.5, 1.75, 2.0])
    np.testing.assert_allclose(b["times"], t, rtol=0, atol=1e-15)
    want = np.concatenate([[0.0], 0.01 * np.sin(2 * np.pi * t[1:5]), 0.02 * np.sin(4 * np.pi * t[5:] + np.pi / 2)])
    np.testing.assert_allclose(b["values"], want, rtol=0, atol=1e-17)
    for k, (bad, msg) in enumerate([({"phase": None}, '"phase" was not specified'),
                                    ({"period": [1.0]}, "must be of the same size"),
                                    ({"generate_type": "square"}, "not recognised")]):
        bb = dict(bc, **bad)
        bb = {kk: v for kk, v in bb.items() if v is not None}
        path = neon_files.write_cube_simulation(tmp_path / f"b{k}", 2, NEOHOOKE, {"name": "neohooke"},
                                                neon_files.rollers([bb]), time=time)
        with pytest.raises(ValueError, match=msg):
            ingest.read(path)
    bb = {k: v for k, v in bc.items() if k != "generate_type"}
    path = neon_files.write_cube_simulation(tmp_path / "c", 2, NEOHOOKE, {"name": "neohooke"},
                                            neon_files.rollers([bb]), time=time)
    with pytest.raises(ValueError, match='Neither "time" nor "generate_type"'):
        ingest.read(path)


def sinusoid_problem(tmp_path, n=2):
    """Three quarters of a compression-first cycle of u_z on z1 (generated history, step 0.25)."""
    time = {"period": 0.75, "increments": {"initial": 0.25, "minimum": 0.01, "maximum": 0.25, "adaptive": False}}
    bcs = [{"name": f"{a}0", "type": "displacement", "time": [0.0, 0.75], a: [0.0, 0.0]} for a in "xyz"]
    bcs.append(_sinusoid_bc(z=[-0.05], period=[1.0], number_of_cycles=[0.75], phase=[0.0]))
    return ingest.read(neon_files.write_cube_simulation(tmp_path, n, NEOHOOKE, {"name": "neohooke"}, bcs, time=time))


def test_generated_history_drives_the_newton_solve(tmp_path):
    p = sinusoid_problem(tmp_path)
    code, u, recs, _ = oracle_solve(p)
    assert code == orc.OK
    assert sorted({round(r["step_time"], 12) for r in recs}) == [0.25, 0.5, 0.75]
    # u_z = -0.05 sin(3 pi / 2) X_z = 0.05 X_z at the end, homogeneous
    np.testing.assert_allclose(u.reshape(-1, 3)[:, 2], 0.05 * p.X[:, 2], rtol=0, atol=1e-12)


def test_viscoplastic_damage_cube_example():
    # the reference's viscoplastic damage cube (examples/solid_mechanics/viscoplastic_damage/cube):
    # one hex8, symmetry rollers, the top pulled to 1.8 % over 10 ms in 1 ms increments, the
    # "isotropic_chaboche" J2 model with the material's six damage constants
    p = ingest.read(DAMAGE_CUBE)
    assert p.model == _capi.MODEL_J2_DAMAGE
    assert p.damage == problems.DAMAGE_CONSTANTS
    E, nu = 134.0e3, 0.3
    np.testing.assert_allclose(p.params[0], [E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu)), 85.0, 0.0], rtol=1e-15)
    assert sorted(b["name"] for b in p.boundaries) == ["Xsym", "Ysym", "ZLoad", "Zsym"]
    # PaStiX maps to the device Krylov solvers at a relative 1e-12
    assert p.simulation["linear_solver"]["tolerance"] == 1e-12
    # the oracle (BiCGSTAB on the unsymmetric tangent) runs it to the end with damage growing
    code, u, recs, m = oracle_solve(p)
    assert code == orc.OK
    assert max(r["step_time"] for r in recs) == pytest.approx(0.01)
    d = m.damage_state()["damage"]
    assert (d > 0).all() and (d < 1).all()
    # the same problem as problems.damage_cube(1): the top face's displacement is 0.018
    top = np.nonzero(p.X[:, 2] == p.X[:, 2].max())[0]
    np.testing.assert_allclose(u.reshape(-1, 3)[top, 2], 0.018, rtol=1e-14)
