"""The symmetric sliced copy of the tangent (diagonal and upper blocks only, the lower ones read
as transposes of their upper partners) through the C ABI.

It stands in for the full copy only where the tangent is symmetric: every model that neon
declares symmetric (constitutive_model::is_symmetric) assembles a tangent symmetric to rounding
in both tangent modes (checked here against its transpose, 1e-13 of max |K|), and the J2 damage
model, the one unsymmetric tangent, never gets the copy.  Products match scipy's product of the
downloaded matrix on meshes of unusual shape (one slice, a last slice with two nodes, one cell
thick, long rows), FP64 to 1e-13 and FP32 to 16 FP32 ulps of |K||x| per row; solves and Newton
runs with the copy agree with the full copy's (same Newton iterations, u to 1e-10 of max |u|;
Krylov counts within 2 % in FP64 and 15 % in the mixed solver, whose FP32 recurrence is
sensitive to the order of the row sums)."""
import numpy as np
import pytest
import scipy.sparse as sp

import neon_amd
from neon_amd import _capi, problems

import test_gpu_krylov as tgk
import test_gpu_models as tgm

pytestmark = pytest.mark.gpu

FULL = neon_amd.SPMV_SLICED | neon_amd.SPMV_NONTEMPORAL
SYMMETRIC = FULL | neon_amd.SPMV_SYMMETRIC


def _asymmetry(mg):
    rp, ci = mg.pattern()
    val, _ = mg.download(internal_force=False)
    A = sp.csr_matrix((val, ci, rp))
    return abs(A - A.T).max() / abs(A).max()


SYMMETRIC_MODELS = [neon_amd.MODEL_NEOHOOKE, neon_amd.MODEL_J2] + tgm.MODELS


@pytest.mark.parametrize("tangent", [neon_amd.TANGENT_REFERENCE, neon_amd.TANGENT_CONSISTENT])
@pytest.mark.parametrize("model", SYMMETRIC_MODELS)
def test_symmetric_models_assemble_symmetric_tangents(ctx, model, tangent):
    print(f"model {model:#x} tangent {tangent}", flush=True)
    X, conn = tgm._block()
    if model == neon_amd.MODEL_J2:
        n = len(conn)
        prm = np.stack([np.full(n, 3.0), np.full(n, 1.0), np.full(n, 2e-3), np.full(n, 0.05)], axis=1)
        scale = 0.01
    else:
        prm = tgm._params(model, len(conn)) if model != neon_amd.MODEL_NEOHOOKE else tgm._params(
            neon_amd.MODEL_LINEAR_ELASTIC, len(conn))
        scale = 0.15
    mg = neon_amd.DeviceMesh(X, conn, prm, model=model, tangent=tangent, context=ctx)
    assert mg.symmetric_info()["usable"]
    mg.set_displacement(tgm._deformation(X, scale))
    mg.update_internal_variables()
    if model == neon_amd.MODEL_J2:
        assert (mg.qp_state()["ebar"] > 0).sum() > 100  # the plastic tangent is in play
    for precision in (neon_amd.PRECISION_FP64, neon_amd.PRECISION_MIXED):
        mg.assemble(precision)
        asym = _asymmetry(mg)
        print("precision", precision, "max |K - K^T| / max |K|", asym)
        assert asym <= 1e-13, asym


def test_symmetric_copy_only_where_it_applies(ctx):
    # a structured block: 13 offsets per slice (the neighbours above and below in x, in the
    # next and previous row and plane); node numbers shuffled: far more than 40 per slice; the J2
    # damage model (unsymmetric tangent) on the structured block: the layout is there but the
    # copy is not used
    p = problems.block_rollers(6)
    mg = neon_amd.DeviceMesh(p.X, p.conn, p.params, context=ctx)
    info = mg.symmetric_info()
    assert info == {"usable": True, "offsets": 13, "slots": info["slots"]}, info
    assert info["slots"] % 32 == 0 and info["slots"] >= (p.n_nodes + 63) // 64
    perm = np.random.default_rng(3).permutation(p.n_nodes)
    inv = np.argsort(perm)
    shuffled = neon_amd.DeviceMesh(p.X[perm], inv[p.conn.reshape(-1, 8)].astype(np.int32), p.params, context=ctx)
    assert shuffled.symmetric_info() == {"usable": False, "offsets": 0, "slots": 0}
    damage = neon_amd.DeviceMesh(p.X, p.conn, p.params, model=neon_amd.MODEL_J2_DAMAGE, context=ctx)
    info = damage.symmetric_info()
    assert not info["usable"] and info["offsets"] == 13, info


# (cells): one slice (27 nodes); 66 nodes (the last slice holds two; rows of 2 nodes, so the
# neighbours one row up share offsets with the next node in the row); a slab one cell thick;
# rows of 121 nodes (slices straddle rows); the 12^3 block; plates with a hole, whose node
# numbering skips the hole, so that slices near it have 16 and 28 offsets of their own (the
# second plate's copy holds 0.74 of the full copy's steps, near the 3/4 limit)
SHAPES = [dict(n=2), dict(n=4, dims=(1, 2, 10)), dict(n=3, dims=(9, 5, 1)), dict(n=12, dims=(120, 3, 3)),
          dict(n=12), dict(plate=(12, 6, 3, 1.5)), dict(plate=(24, 12, 6, 3.0))]


def _mesh(ctx, spec):
    if "plate" in spec:
        nx, ny, nz, hole = spec["plate"]
        p = problems.notched_plate_j2(nx, ny, nz, steps=2, hole_cells=hole)
    else:
        p = problems.block_rollers(**spec)
    mg = neon_amd.DeviceMesh(p.X, p.conn, p.params, context=ctx)
    return p, mg


@pytest.mark.parametrize("spec", SHAPES)
@pytest.mark.parametrize("precision", [neon_amd.PRECISION_FP64, neon_amd.PRECISION_MIXED])
def test_symmetric_products_match_the_matrix(ctx, precision, spec):
    p, mg = _mesh(ctx, spec)
    info = mg.symmetric_info()
    print(spec, "nodes", p.n_nodes, info, flush=True)
    assert info["usable"]
    mg.set_displacement(tgk._smooth_state(p, 0.05, 7))
    mg.update_internal_variables()
    mg.assemble(precision)
    mg.enforce_dirichlet(p.fixed_dofs(), None)
    rp, ci = mg.pattern()
    val, _ = mg.download(internal_force=False)
    A = sp.csr_matrix((val, ci, rp))
    x = np.random.default_rng(5).standard_normal(p.n_dofs)
    fp32 = precision == neon_amd.PRECISION_MIXED
    xr = x.astype(np.float32).astype(np.float64) if fp32 else x
    ref = A @ xr
    scale = abs(A) @ np.abs(xr)
    tol = 16 * np.finfo(np.float32).eps if fp32 else 1e-13
    lib = _capi.measure_lib()
    for extra in (0, _capi.MEASURE_DOT):
        # 0: the grid of the schedule; 8 and 16: workgroups striding over its wave slots
        for blocks in (0, 8, 16):
            y = np.zeros(p.n_dofs)
            ctx.check(lib.neon_mesh_spmv_apply(ctx.handle, precision, 32, SYMMETRIC | extra, blocks,
                                               _capi.f64(np.ascontiguousarray(x)), _capi.f64(y)))
            err = (np.abs(y - ref) / scale).max()
            print("flags", SYMMETRIC | extra, "blocks", blocks, "max error / (|K||x|)", err)
            assert err <= tol, (extra, blocks, err)


def test_mixed_solve_on_the_scaled_copy(ctx):
    # on the symmetric copy the mixed solver's FP32 recurrence runs on the Jacobi-scaled tangent
    # D^-1/2 K D^-1/2 (a copy of its own) while its FP64 residuals use the unscaled FP32 values:
    # the solution meets the tolerance against the downloaded FP32 matrix, agrees with the full
    # copy's (unscaled Jacobi-PCG) solve, and the products of the unscaled copy afterwards still
    # match the matrix
    p = problems.block_rollers(12)
    mg = neon_amd.DeviceMesh(p.X, p.conn, p.params, context=ctx)
    mg.set_displacement(tgk._smooth_state(p, 0.05, 7))
    mg.update_internal_variables()
    mg.assemble(neon_amd.PRECISION_MIXED)
    _, f = mg.download(stiffness=False)
    b = -f.copy()
    mg.enforce_dirichlet(p.fixed_dofs(), b)
    rp, ci = mg.pattern()
    val, _ = mg.download(internal_force=False)
    A = sp.csr_matrix((val, ci, rp))
    sols = {}
    for name, flags in (("full", FULL), ("symmetric", SYMMETRIC)):
        mg.spmv_config(32, flags)
        x, its, _ = mg.solve(b, criterion=neon_amd.STOP_RELATIVE, tol=1e-10)
        res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        print(name, "iterations", its, "relative residual against the FP32 matrix", res, flush=True)
        assert res <= 1e-10
        sols[name] = (x, its)
    (x0, i0), (x1, i1) = sols["full"], sols["symmetric"]
    assert abs(i1 - i0) <= max(8, 0.15 * i0)
    assert np.linalg.norm(x1 - x0) <= 1e-8 * np.linalg.norm(x0)
    y = np.zeros(p.n_dofs)
    xr = np.random.default_rng(5).standard_normal(p.n_dofs)
    xr32 = xr.astype(np.float32).astype(np.float64)
    ctx.check(_capi.measure_lib().neon_mesh_spmv_apply(ctx.handle, neon_amd.PRECISION_MIXED, 32, SYMMETRIC, 0,
                                                       _capi.f64(np.ascontiguousarray(xr)), _capi.f64(y)))
    err = (np.abs(y - A @ xr32) / (abs(A) @ np.abs(xr32))).max()
    print("unscaled product after the scaled solve: max error / (|K||x|)", err)
    assert err <= 16 * np.finfo(np.float32).eps
    mg.spmv_config()


def test_mixed_solve_without_a_positive_diagonal(ctx):
    # negated material parameters: at u = 0 the tangent is negative definite with a negative
    # diagonal, on which Jacobi-PCG still converges (M^-1 K is similar to a positive definite
    # matrix) but the scaled recurrence, which needs sqrt(M^-1), does not apply: the solver keeps
    # the unscaled recurrence on the symmetric copy (a scaled run would turn the square roots
    # into NaN), meets the tolerance against the downloaded FP32 matrix and agrees with the full
    # copy's solve
    p = problems.block_rollers(8)
    mg = neon_amd.DeviceMesh(p.X, p.conn, -p.params, context=ctx)
    assert mg.symmetric_info()["usable"]
    mg.set_displacement(np.zeros(p.n_dofs))
    mg.update_internal_variables()
    mg.assemble(neon_amd.PRECISION_MIXED)
    b = np.random.default_rng(11).standard_normal(p.n_dofs)
    mg.enforce_dirichlet(p.fixed_dofs(), b)
    rp, ci = mg.pattern()
    val, _ = mg.download(internal_force=False)
    A = sp.csr_matrix((val, ci, rp))
    assert (A.diagonal() < 0).all()
    sols = []
    for flags in (FULL, SYMMETRIC):
        mg.spmv_config(32, flags)
        x, its, _ = mg.solve(b, criterion=neon_amd.STOP_RELATIVE, tol=1e-10)
        res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        print("flags", flags, "iterations", its, "relative residual", res, flush=True)
        assert np.isfinite(x).all() and res <= 1e-10
        sols.append((x, its))
    (x0, i0), (x1, i1) = sols
    assert abs(i1 - i0) <= max(8, 0.15 * i0)
    assert np.linalg.norm(x1 - x0) <= 1e-8 * np.linalg.norm(x0)
    mg.spmv_config()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_symmetric_copy_newton_matches_the_full_copy(ctx, precision):
    # the 12^3 block, 2 load steps: the same Newton steps and iterations with either copy, the
    # displacement to 1e-10 of max |u|, Krylov counts within 2 % (FP64) or 15 % (mixed)
    p = problems.block_rollers(12, steps=2, stretch=0.04, precision=precision)
    runs = []
    for flags in (FULL, SYMMETRIC):
        mg = neon_amd.DeviceMesh(p.X, p.conn, p.params, context=ctx)
        mg.spmv_config(32, flags)
        sm = neon_amd.static_matrix(mg, p.simulation, p.boundaries)
        u = sm.solve()
        runs.append((u, sm.records))
        print("flags", flags, "newton", len(sm.records), "krylov", [r["solver_iterations"] for r in sm.records],
              flush=True)
    (u0, r0), (u1, r1) = runs
    key = lambda rr: [(round(r["step_time"], 12), r["iteration"], r["status"]) for r in rr]  # noqa: E731
    assert key(r0) == key(r1)
    slack = 0.02 if precision == "fp64" else 0.15
    for a, b in zip(r0, r1):
        assert abs(a["solver_iterations"] - b["solver_iterations"]) <= max(2, slack * a["solver_iterations"])
    assert np.abs(u1 - u0).max() <= 1e-10 * np.abs(u0).max()


@pytest.mark.parametrize("hole", [0.0, 1.5])
@pytest.mark.parametrize("tangent", [neon_amd.TANGENT_REFERENCE, neon_amd.TANGENT_CONSISTENT])
def test_symmetric_copy_j2_plate_matches_the_full_copy(ctx, tangent, hole):
    # a J2 plate pulled to 0.5 %, well into yield: 12 x 6 x 6 cells without the hole (a
    # structured pattern) in 4 load steps, and the parity tests' 12 x 6 x 3 plate with the hole
    # (slices near it have offsets of their own) on its 20-step schedule and norms (with fewer
    # steps or neon's default norms, neon's mode exhausts its cut-backs): with either copy the
    # same Newton steps.  In the consistent mode Newton counts per step within one (as in the
    # parity tests: a step's last residual can land within roundoff of the tolerance), the
    # displacement to 1e-10 of max |u| and the same plastic zone.  Neon's mode updates the history in place at every
    # Newton iterate, so its converged state follows the roundoff of the iterates: on the holed
    # plate the oracle's own runs differ by up to 2.8e-4 of max |u| between thread counts
    # (tests/test_gpu_parity.py measures that envelope); the two copies, whose products differ
    # by rounding, are held to 1e-3 there and to 1 % of the quadrature points in the plastic zone
    if hole:
        p = problems.notched_plate_j2(12, 6, 3, steps=20, hole_cells=hole, residual_tolerance=1e-8,
                                      displacement_tolerance=1e-8, norm_mode=1)
    else:
        p = problems.notched_plate_j2(12, 6, 6, steps=4, hole_cells=0.0)
    runs = []
    for flags in (FULL, SYMMETRIC):
        mg = neon_amd.DeviceMesh(p.X, p.conn, p.params, model=neon_amd.MODEL_J2, tangent=tangent, context=ctx)
        assert mg.symmetric_info()["usable"]
        mg.spmv_config(32, flags)
        sm = neon_amd.static_matrix(mg, p.simulation, p.boundaries)
        u = sm.solve()
        runs.append((u, sm.records, mg.qp_state()["ebar"]))
        print("flags", flags, "newton", len(sm.records), "krylov", [r["solver_iterations"] for r in sm.records],
              "yielded", int((runs[-1][2] > 0).sum()), flush=True)
    (u0, r0, e0), (u1, r1, e1) = runs
    steps = lambda rr: sorted({round(r["step_time"], 12) for r in rr})  # noqa: E731
    assert steps(r0) == steps(r1) and all(r["status"] == 0 for r in r0 + r1)
    assert (e0 > 0).sum() > 0.25 * e0.size
    err = np.abs(u1 - u0).max() / np.abs(u0).max()
    print("u difference", err, "plastic zone mismatch", int(((e0 > 0) != (e1 > 0)).sum()))
    if tangent == neon_amd.TANGENT_CONSISTENT:
        for t in steps(r0):
            n0 = sum(round(r["step_time"], 12) == t for r in r0)
            n1 = sum(round(r["step_time"], 12) == t for r in r1)
            assert abs(n0 - n1) <= 1, (t, n0, n1)
        assert np.array_equal(e0 > 0, e1 > 0)
        assert err <= 1e-10
    else:
        assert ((e0 > 0) != (e1 > 0)).sum() <= 0.01 * e0.size
        assert err <= 1e-3
