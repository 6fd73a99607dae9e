"""The constitutive models beyond neo-Hooke and J2 on the HIP path (isotropic linear elasticity,
the Gaussian and Langevin affine and the non-affine microsphere models with the BO21 and BO33
rules), through the C ABI, against the CPU oracle (pinned in test_microsphere.py).

Tolerances as for the other models: per-QP stress and tangent to 1e-12 of their maximum, tangent
entries to 1e-12 of their row maximum, f_int to 1e-12 of its maximum; Newton runs with equal
step and iteration sequences and u to 1e-10 of max |u| (mixed precision: 1e-9)."""
import numpy as np
import pytest

import neon_amd
from neon_amd import ingest
from oracle import oracle as orc

import neon_files
import test_gpu_ingest as tgi
import test_ingest as ti

pytestmark = pytest.mark.gpu

MODELS = [
    orc.MODEL_LINEAR_ELASTIC,
    orc.MODEL_MICROSPHERE_GAUSSIAN,
    orc.MODEL_MICROSPHERE_GAUSSIAN | orc.SPHERE_BO33,
    orc.MODEL_MICROSPHERE_LANGEVIN,
    orc.MODEL_MICROSPHERE_LANGEVIN | orc.SPHERE_BO33,
    orc.MODEL_MICROSPHERE_NONAFFINE,
    orc.MODEL_MICROSPHERE_NONAFFINE | orc.SPHERE_BO33,
]


def _block(n=4, seed=6):
    X, conn, _ = neon_files.cube_mesh(n)
    rng = np.random.default_rng(seed)
    inner = np.all((X > 0) & (X < 1), axis=1)
    X = X.copy()
    X[inner] += rng.uniform(-0.12 / n, 0.12 / n, (inner.sum(), 3))
    return X, conn


def _params(model, n_elem, seed=7):
    rng = np.random.default_rng(seed)
    K = 3.0 * rng.uniform(0.8, 1.25, n_elem)
    G = rng.uniform(0.8, 1.25, n_elem)
    if (model & 0xFF) == orc.MODEL_LINEAR_ELASTIC:
        return np.stack([K, G, np.zeros(n_elem), np.zeros(n_elem)], axis=1)
    N = rng.uniform(6.0, 12.0, n_elem)
    p = rng.uniform(1.2, 2.5, n_elem) if (model & 0xFF) == orc.MODEL_MICROSPHERE_NONAFFINE else np.zeros(n_elem)
    return np.stack([K, G, N, p], axis=1)


def _deformation(X, scale):
    u = np.zeros_like(X)
    u[:, 2] = scale * X[:, 2] * (1 + 0.2 * np.sin(3 * X[:, 0]))
    u[:, 0] = -0.3 * scale * X[:, 0] * (1 + 0.1 * np.cos(2 * X[:, 1]))
    u[:, 1] = 0.05 * scale * np.sin(2 * X[:, 2])
    return u.reshape(-1)


def _row_scaled(val, val0, rp):
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    rowmax = np.maximum.reduceat(np.abs(val0), rp[:-1])
    return (np.abs(val - val0) / rowmax[rows]).max()


@pytest.mark.parametrize("tangent", [neon_amd.TANGENT_REFERENCE, neon_amd.TANGENT_CONSISTENT])
@pytest.mark.parametrize("model", MODELS)
def test_state_and_assembly_parity(ctx, model, tangent):
    print(f"model {model:#x} tangent {tangent}", flush=True)
    X, conn = _block()
    prm = _params(model, len(conn))
    m_gpu = neon_amd.DeviceMesh(X, conn, prm, model=model, tangent=tangent, context=ctx)
    m_ref = orc.Model(X, conn, prm, model=model, tangent=tangent)
    u = _deformation(X, 0.15)
    m_gpu.set_displacement(u)
    m_gpu.update_internal_variables()
    assert m_ref.update(u) == orc.OK
    sg, sr = m_gpu.qp_state(), m_ref.qp_state()
    keys = ("sigma", "D", "detF") + (("vm",) if model == orc.MODEL_LINEAR_ELASTIC else ())
    for key in keys:
        ref = sr[key]
        np.testing.assert_allclose(sg[key], ref, rtol=0, atol=1e-12 * np.abs(ref).max(), err_msg=key)
    rp, ci = m_ref.pattern()
    val0, f0 = m_ref.assemble(rp, ci)
    m_gpu.assemble(neon_amd.PRECISION_FP64)
    val, f = m_gpu.download()
    assert _row_scaled(val, val0, rp) < 1e-12
    np.testing.assert_allclose(f, f0, rtol=0, atol=1e-12 * np.abs(f0).max())
    # the FP32 tangent of the mixed path: the FP64 values rounded once
    m_gpu.assemble(neon_amd.PRECISION_MIXED)
    val32, _ = m_gpu.download()
    assert _row_scaled(val32, val0, rp) < 2e-7


def _simulation(tmp_path, model, bcs, n=3):
    mat = {"elastic_modulus": 8.0, "poissons_ratio": 0.35, "segments_per_chain": 9.0,
           "nonaffine_stretch_parameter": 1.6}
    base, rule = model & 0xFF, "BO33" if model & orc.SPHERE_BO33 else "BO21"
    constitutive = {
        orc.MODEL_LINEAR_ELASTIC: {"name": "isotropic_linear_elasticity"},
        orc.MODEL_MICROSPHERE_GAUSSIAN: {"name": "microsphere", "type": "affine", "statistics": "gaussian",
                                         "quadrature": rule},
        orc.MODEL_MICROSPHERE_LANGEVIN: {"name": "microsphere", "type": "affine", "statistics": "langevin",
                                         "quadrature": rule},
        orc.MODEL_MICROSPHERE_NONAFFINE: {"name": "microsphere", "type": "nonaffine", "quadrature": rule},
    }[base]
    p = ingest.read(neon_files.write_cube_simulation(tmp_path, n, mat, constitutive, bcs))
    assert p.model == model
    return p


@pytest.mark.parametrize("model", MODELS)
def test_newton_uniaxial_stretch(ctx, tmp_path, model):
    """Rollers and a prescribed top displacement: the device Newton repeats the oracle's steps
    and iterations and lands on its displacement, the homogeneous uniaxial state.  The non-affine
    model is stretched less: with neon's tangent (which lacks the -2/3 (dev tau (x) 1 + ...) term,
    large for its tube stresses) its Newton iterations converge slowly, and at 10 % stretch the
    oracle's load stepping gives up."""
    print(f"model {model:#x}", flush=True)
    stretch = 0.05 if (model & 0xFF) == orc.MODEL_MICROSPHERE_NONAFFINE else 0.2
    p = _simulation(tmp_path, model, neon_files.rollers(
        [{"name": "z1", "type": "displacement", "time": [0.0, 1.0], "z": [0.0, stretch]}]))
    u, recs = tgi.gpu_solve(p, ctx)
    code, u0, recs0, _ = ti.oracle_solve(p)
    assert code == orc.OK
    tgi._same_newton(recs, recs0)
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-10 * np.abs(u0).max())
    U = u.reshape(-1, 3)
    np.testing.assert_allclose(U[:, 2], stretch * p.X[:, 2], rtol=0, atol=1e-9)


@pytest.mark.parametrize("model", [orc.MODEL_MICROSPHERE_LANGEVIN, orc.MODEL_MICROSPHERE_NONAFFINE | orc.SPHERE_BO33])
def test_newton_under_traction_mixed_precision(ctx, tmp_path, model):
    """A traction on z1 and the mixed-precision path: the converged displacement of the FP64
    oracle."""
    p = _simulation(tmp_path, model, neon_files.rollers(
        [{"name": "z1", "type": "traction", "time": [0.0, 1.0], "z": [0.0, 0.4]}]), n=4)
    u, recs = tgi.gpu_solve(p, ctx, precision="mixed")
    code, u0, recs0, _ = ti.oracle_solve(p)
    assert code == orc.OK
    assert [round(r["step_time"], 12) for r in recs if r["iteration"] == 0] == \
        [round(r["step_time"], 12) for r in recs0 if r["iteration"] == 0]
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-9 * np.abs(u0).max())


def test_linear_elastic_output_fields(ctx, tmp_path):
    """The output files of a linear-elasticity run: the written linearised strain and von Mises
    stress (device state, node-averaged) equal the homogeneous uniaxial closed form."""
    import xml.etree.ElementTree as ET

    from neon_amd import output

    p = _simulation(tmp_path, orc.MODEL_LINEAR_ELASTIC, neon_files.rollers(
        [{"name": "z1", "type": "displacement", "time": [0.0, 1.0], "z": [0.0, 0.001]}]))
    p.visualisation = {"fields": ["linearised_strain", "von_mises_stress", "displacement"]}
    p.output_name = "cube"
    out = tmp_path / "out"
    u, _, w = output.solve_with_output(p, ctx, str(out))
    K, G = p.params[0, :2]
    E = 9 * K * G / (3 * K + G)
    nu = (3 * K - 2 * G) / (2 * (3 * K + G))
    piece = ET.parse(out / w.written[-1]).getroot().find("UnstructuredGrid").find("Piece")
    arrays = {a.get("Name"): np.array(a.text.split(), dtype=float).reshape(p.n_nodes, -1)
              for a in piece.find("PointData").findall("DataArray")}
    np.testing.assert_allclose(arrays["displacement"].reshape(-1), u, rtol=0, atol=0)
    np.testing.assert_allclose(arrays["von_mises_stress"][:, 0], E * 1e-3, rtol=1e-8)
    np.testing.assert_allclose(arrays["linearised_strain"][:, 8], 1e-3, rtol=1e-8)
    np.testing.assert_allclose(arrays["linearised_strain"][:, 0], -nu * 1e-3, rtol=1e-8)
