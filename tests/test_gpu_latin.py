"""The LATIN iterations on the HIP path (neon_latin_solve: the device LATIN force,
latin_force_kernel, in the Newton driver) against the CPU oracle's restatement (pinned in
test_latin.py): equal step and iteration sequences, force norms to 1e-8 relative (converged
ones to 1e-12 of the run's largest) and u to 1e-10
of max |u|, for the reference's J2 cube and the Langevin microsphere (general D), at s = 0
(which converges) and on both at s = 0.5 (neon's iterations stall and the load step is cut
back until the minimum step is reached: the same cut-back sequence, force norms and last
iterate)."""
import copy

import numpy as np
import pytest

import neon_amd
from neon_amd import ingest
from oracle import oracle as orc

import neon_files
import test_gpu_models as tgm
import test_ingest as ti

pytestmark = pytest.mark.gpu


def _run(p, ctx, s, sim=None, precision="fp64"):
    sim = copy.deepcopy(sim or p.simulation)
    sim["nonlinear_options"]["latin_search_direction"] = s
    sim["linear_solver"] = dict(sim["linear_solver"], backend={"type": "hip", "precision": precision})
    mesh = neon_amd.DeviceMesh(p.X, p.conn, p.params, model=p.model, tangent=neon_amd.TANGENT_REFERENCE, context=ctx)
    lm = neon_amd.latin_matrix(mesh, sim, p.boundaries, loads=p.loads)
    try:
        u, err = lm.solve(), None
    except neon_amd.NeonError as e:
        u, err = None, e
    m = orc.Model(p.X, p.conn, p.params, model=p.model, tangent=orc.TANGENT_REFERENCE)
    code, u0, recs0 = m.newton(orc.newton_options(sim, latin=s), p.boundaries, loads=p.loads)
    return u, err, lm.records, mesh, code, u0, recs0


def _same(recs, recs0):
    steps = lambda rr: [(round(r["step_time"], 12), r["iteration"], r["status"]) for r in rr]  # noqa: E731
    assert steps(recs) == steps(recs0)
    # a converged residual is a difference of forces of the first iteration's size, so it is
    # compared to 1e-12 of the largest norm of the run (rounding), the others to 1e-8
    f, f0 = np.array([r["force_norm"] for r in recs]), np.array([r["force_norm"] for r in recs0])
    np.testing.assert_allclose(f, f0, rtol=1e-8, atol=1e-12 * f0.max())


@pytest.mark.parametrize("case", ["j2_cube", "langevin"])
def test_latin_zero_direction(ctx, tmp_path, case):
    print(f"latin s=0 {case}", flush=True)
    if case == "j2_cube":
        p = ingest.read(ti.J2_CUBE)
    else:
        p = tgm._simulation(tmp_path, orc.MODEL_MICROSPHERE_LANGEVIN, neon_files.rollers(
            [{"name": "z1", "type": "displacement", "time": [0.0, 1.0], "z": [0.0, 0.2]}]))
    u, err, recs, mesh, code, u0, recs0 = _run(p, ctx, 0.0)
    assert err is None and code == orc.OK
    _same(recs, recs0)
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-10 * np.abs(u0).max())


def test_latin_stalling_direction_cut_backs(ctx):
    print("latin s=0.5 j2_cube", flush=True)
    p = ingest.read(ti.J2_CUBE)
    u, err, recs, mesh, code, u0, recs0 = _run(p, ctx, 0.5)
    assert err is not None and code != orc.OK
    _same(recs, recs0)
    # the last iterate of the failed attempt (neither driver resets u on the fatal exit)
    np.testing.assert_allclose(mesh.displacement(), u0, rtol=0, atol=1e-8 * max(np.abs(u0).max(), 1e-300))


def test_latin_general_d_nonzero_direction(ctx, tmp_path):
    # the Langevin microsphere's general (21-coefficient) D enters the LATIN force through
    # s D (eps - eps_old) only when s != 0: at s = 0.5 neon's iterations stall on this model as
    # on the J2 cube (the oracle cuts the step back to the minimum and gives up after 50
    # iterations), so the device must repeat every iteration's force norm (to 1e-8), the
    # cut-back sequence and the last iterate
    print("latin s=0.5 langevin", flush=True)
    p = tgm._simulation(tmp_path, orc.MODEL_MICROSPHERE_LANGEVIN, neon_files.rollers(
        [{"name": "z1", "type": "displacement", "time": [0.0, 1.0], "z": [0.0, 0.2]}]))
    u, err, recs, mesh, code, u0, recs0 = _run(p, ctx, 0.5)
    assert err is not None and code != orc.OK
    assert len(recs0) > 10 and max(r["force_norm"] for r in recs0) > 0
    _same(recs, recs0)
    np.testing.assert_allclose(mesh.displacement(), u0, rtol=0, atol=1e-8 * max(np.abs(u0).max(), 1e-300))


def test_latin_mixed_precision(ctx, tmp_path):
    """The LATIN right-hand side through the mixed-precision PCG: the converged displacement of
    the FP64 oracle to 1e-9 of max |u| and the same load steps."""
    p = tgm._simulation(tmp_path, orc.MODEL_MICROSPHERE_GAUSSIAN, neon_files.rollers(
        [{"name": "z1", "type": "displacement", "time": [0.0, 1.0], "z": [0.0, 0.2]}]), n=4)
    u, err, recs, mesh, code, u0, recs0 = _run(p, ctx, 0.0, precision="mixed")
    assert err is None and code == orc.OK
    assert [round(r["step_time"], 12) for r in recs if r["iteration"] == 0] == \
        [round(r["step_time"], 12) for r in recs0 if r["iteration"] == 0]
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-9 * np.abs(u0).max())
