"""HIP BiCGSTAB and mixed-precision PCG against the CPU FP64 oracle (through the C ABI).

Mixed precision stops on the true FP64 residual of the FP32-valued tangent, so its solution
is compared with the oracle's FP64 solve of that same FP32-rounded matrix (1e-8 relative to
max|x|), and the true residual is checked directly (<= tol ||b||).  Newton in mixed precision
converges to the FP64 f_int root; its displacement is compared with the FP64 oracle run
(error stated per test)."""
import numpy as np
import pytest
import scipy.sparse as sp

import neon_amd
from neon_amd import problems
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _smooth_state(p, scale, seed):
    X = p.X
    u = np.zeros_like(X)
    u[:, 2] = scale * X[:, 2] * (1 + 0.2 * np.sin(3 * X[:, 0]))
    u[:, 0] = -0.3 * scale * X[:, 0] * (1 + 0.1 * np.cos(2 * X[:, 1]))
    return u.reshape(-1) + 0.01 * scale * np.random.default_rng(seed).normal(size=p.n_dofs)


def _system(p, scale=0.01, seed=3):
    m = orc.Model(p.X, p.conn, p.params, model=p.model)
    u = _smooth_state(p, scale, seed)
    assert m.update(u) == orc.OK
    rp, ci = m.pattern()
    val, f = m.assemble(rp, ci)
    b = -f + 1e-3 * np.random.default_rng(seed + 1).normal(size=p.n_dofs)
    orc.enforce_dirichlet(rp, ci, val, b, p.fixed_dofs())
    return rp, ci, val, b


@pytest.mark.parametrize("criterion", [neon_amd.STOP_RELATIVE, neon_amd.STOP_ABSOLUTE_PRECONDITIONED])
def test_bicgstab_host_csr_parity(ctx, criterion):
    p = problems.cube_clamped(6)
    rp, ci, val, b = _system(p)
    tol = 1e-10 if criterion == neon_amd.STOP_RELATIVE else 1e-12
    ref = orc.bicgstab_eigen if criterion == neon_amd.STOP_RELATIVE else orc.bicgstab_cuda
    code, x0, its0, _ = ref(rp, ci, val, b, tol, 20000)
    assert code == orc.OK
    s = neon_amd.biconjugate_gradient_stabilised_hip(tol, 20000, criterion=criterion, context=ctx)
    x = np.zeros_like(b)
    s.solve((rp, ci, val), x, b)
    # BiCGSTAB counts follow roundoff closely: the oracle's own count on this system moves
    # between 89 and 103 (Eigen rule) with its thread count (the order of its parallel dot
    # products), so the count is held within 15 % of the oracle's range over several thread
    # counts.  The CUDA rule's test value sqrt(r0 . r) is not a norm and its count moves more
    # (39 to 50 between hosts), so only the attained residual is compared there.
    if criterion == neon_amd.STOP_RELATIVE:
        counts = [its0] + [ref(rp, ci, val, b, tol, 20000, threads=t)[2] for t in (1, 2, 3, 4, 8)]
        assert 0.85 * min(counts) <= s.iterations <= 1.15 * max(counts), (s.iterations, counts)
    A = sp.csr_matrix((val, ci, rp))
    res, res0 = np.linalg.norm(b - A @ x), np.linalg.norm(b - A @ x0)
    if criterion == neon_amd.STOP_RELATIVE:
        assert res <= 1.01e-10 * np.linalg.norm(b)
        np.testing.assert_allclose(x, x0, rtol=0, atol=1e-7 * np.abs(x0).max())
    else:
        # the CUDA rule stops on sqrt(r0 . r), which is not a residual norm (it can cancel
        # early), so only the attained residuals are compared, against the range the oracle's
        # own residual spans over its thread counts (8.5e-8 to 1.2e-5 on this system: the
        # cancellation follows the summation order of the dot products)
        res_t = [np.linalg.norm(b - A @ ref(rp, ci, val, b, tol, 20000, threads=t)[1]) for t in (1, 2, 3, 4, 8)]
        assert res <= 10 * max([res0] + res_t) + 1e-14 * np.linalg.norm(b), (res, res0, res_t)


@pytest.mark.parametrize("solver", ["cg", "bicgstab"])
@pytest.mark.parametrize("criterion", [neon_amd.STOP_RELATIVE, neon_amd.STOP_ABSOLUTE_PRECONDITIONED])
def test_krylov_iterates_follow_the_oracle(ctx, solver, criterion):
    # the solvers' k-th iterate against the oracle's (same algorithm, same stopping rule, a
    # tolerance that never stops them), before roundoff has had time to grow: the oracle's own
    # iterates move by 2e-15 (BiCGSTAB at k = 10) and 1e-15 (CG at k = 40) of max|x| between
    # its thread counts, so a wrong coefficient, sign or update order shows at once
    p = problems.cube_clamped(6)
    rp, ci, val, b = _system(p)
    if solver == "cg":
        ref = orc.cg_eigen if criterion == neon_amd.STOP_RELATIVE else orc.cg_cuda
        cls, ks, bound = neon_amd.conjugate_gradient_hip, (5, 10, 40), 1e-13
    else:
        ref = orc.bicgstab_eigen if criterion == neon_amd.STOP_RELATIVE else orc.bicgstab_cuda
        cls, ks, bound = neon_amd.biconjugate_gradient_stabilised_hip, (5, 10), 1e-13
    for k in ks:
        code0, x0, its0, _ = ref(rp, ci, val, b, 1e-30, k)
        s = cls(1e-30, k, criterion=criterion, context=ctx)
        x = np.zeros_like(b)
        with pytest.raises(neon_amd.computational_error):  # the iteration limit is reached
            s.solve((rp, ci, val), x, b)
        err = np.abs(x - x0).max() / np.abs(x0).max()
        print(f"{solver} rule {criterion}: iterate {k}: {s.iterations} vs {its0} iterations, difference {err:.2e}")
        assert s.iterations == its0 and err <= bound, (k, err)


def test_bicgstab_unsymmetric_system(ctx):
    # an unsymmetric diagonally dominant matrix: BiCGSTAB converges where CG would not apply
    rng = np.random.default_rng(9)
    n = 400
    A = sp.random(n, n, density=0.02, random_state=3, format="csr")
    A = A + sp.diags(np.abs(A).sum(axis=1).A1 + 1.0)
    A = A.tocsr()
    A.sort_indices()
    b = rng.normal(size=n)
    s = neon_amd.make_linear_solver({"type": "iterative", "device": "gpu", "tolerance": 1e-12,
                                     "backend": {"type": "hip"}}, is_symmetric=False, context=ctx)
    assert isinstance(s, neon_amd.biconjugate_gradient_stabilised_hip)
    # Eigen's relative rule: the CUDA solver's sqrt(r0.r) test is not a residual norm
    s.criterion = neon_amd.STOP_RELATIVE
    x = np.zeros(n)
    s.solve(A, x, b)
    assert np.linalg.norm(A @ x - b) < 1e-10 * np.linalg.norm(b)


def test_mixed_cg_host_csr(ctx):
    p = problems.cube_clamped(10)
    rp, ci, val, b = _system(p)
    val32 = val.astype(np.float32).astype(np.float64)
    A32 = sp.csr_matrix((val32, ci, rp))
    code, x0, its0, _ = orc.cg_eigen(rp, ci, val32, b, 1e-10, 20000)
    assert code == orc.OK
    s = neon_amd.conjugate_gradient_hip(1e-10, 20000, criterion=neon_amd.STOP_RELATIVE,
                                        precision=neon_amd.PRECISION_MIXED, context=ctx)
    x = np.zeros_like(b)
    s.solve((rp, ci, val), x, b)
    assert np.linalg.norm(b - A32 @ x) <= 1e-10 * np.linalg.norm(b)
    np.testing.assert_allclose(x, x0, rtol=0, atol=1e-8 * np.abs(x0).max())
    # reliable updates keep the Krylov space: iteration count close to the FP64 solver's
    assert s.iterations <= 1.25 * its0 + 5


def test_mixed_device_tangent_solve(ctx):
    p = problems.block_rollers(8)
    mg = neon_amd.DeviceMesh(p.X, p.conn, p.params, context=ctx)
    u = _smooth_state(p, 0.05, 6)
    mg.set_displacement(u)
    mg.update_internal_variables()
    mg.assemble(neon_amd.PRECISION_MIXED)
    _, f = mg.download(stiffness=False)
    rp, ci = mg.pattern()
    b = -f.copy()
    mg.enforce_dirichlet(p.fixed_dofs(), b)
    val32, _ = mg.download(internal_force=False)
    x, its, err = mg.solve(b, criterion=neon_amd.STOP_RELATIVE, tol=1e-10)
    A32 = sp.csr_matrix((val32, ci, rp))
    assert np.linalg.norm(b - A32 @ x) <= 1e-10 * np.linalg.norm(b)
    # FP32 tangent entries equal the FP64 assembly rounded per element contribution
    m = orc.Model(p.X, p.conn, p.params)
    assert m.update(u) == orc.OK
    val64, f64 = m.assemble(rp, ci)
    np.testing.assert_allclose(f, f64, rtol=0, atol=1e-12 * np.abs(f64).max())
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    fixed = set(p.fixed_dofs().tolist())
    free = np.array([r not in fixed and c not in fixed for r, c in zip(rows, ci)])
    rowmax = np.maximum.reduceat(np.abs(val64), rp[:-1])[rows]
    assert (np.abs(val32 - val64)[free] / rowmax[free]).max() < 1e-6


def test_mixed_newton_block(ctx):
    p = problems.block_rollers(6, residual_tolerance=1e-8, displacement_tolerance=1e-8, norm_mode=1,
                               precision="mixed")
    mg = neon_amd.DeviceMesh(p.X, p.conn, p.params, context=ctx)
    sm = neon_amd.static_matrix(mg, p.simulation, p.boundaries)
    u = sm.solve()
    ref = orc.Model(p.X, p.conn, p.params)
    code, u0, recs0 = ref.newton(orc.newton_options(p.simulation), p.boundaries)
    assert code == orc.OK
    assert all(r["status"] == 0 for r in sm.records)
    assert len({r["step"] for r in sm.records}) == len({r["step"] for r in recs0})
    # converged to the same equilibrium: error of the mixed path against the FP64 oracle
    err = np.abs(u - u0).max() / np.abs(u0).max()
    assert err < 1e-7, err
    assert sm.records[-1]["force_norm"] <= 1e-8


@pytest.mark.parametrize("hole, layout", [(1.5, "full"), (1.5, "symmetric"), (0.0, "symmetric")])
def test_mixed_newton_j2_plate(ctx, hole, layout):
    # the J2 plate (consistent tangent) with the inner PCG in FP32: the radial return, f_int,
    # the Newton residual and the reliable updates stay FP64, so every load step converges to the
    # FP64 oracle's equilibrium and the history follows it: the same load steps (Newton counts
    # may differ: the inexact inner solves cost an iteration here and there), u to 1e-9 of
    # max |u| and the accumulated plastic strain to 1e-9 of its maximum (measured 3e-14 and
    # 3e-13), the same plastic zone; on the full sliced copy and on the symmetric one, with the
    # hole (slices near it have offsets of their own) and without
    p = problems.notched_plate_j2(12, 6, 3, steps=20, hole_cells=hole, residual_tolerance=1e-8,
                                  displacement_tolerance=1e-8, norm_mode=1, precision="mixed")
    mg = neon_amd.DeviceMesh(p.X, p.conn, p.params, model=p.model, tangent=neon_amd.TANGENT_CONSISTENT,
                             context=ctx)
    assert mg.symmetric_info()["usable"]
    if layout == "full":
        mg.spmv_config(32, neon_amd.SPMV_SLICED | neon_amd.SPMV_NONTEMPORAL)
    sm = neon_amd.static_matrix(mg, p.simulation, p.boundaries)
    u = sm.solve()
    ref = orc.Model(p.X, p.conn, p.params, model=p.model, tangent=orc.TANGENT_CONSISTENT)
    code, u0, recs0 = ref.newton(orc.newton_options(p.simulation), p.boundaries)
    assert code == orc.OK
    assert all(r["status"] == 0 for r in sm.records)
    assert sorted({round(r["step_time"], 12) for r in sm.records}) == \
        sorted({round(r["step_time"], 12) for r in recs0})
    err = np.abs(u - u0).max() / np.abs(u0).max()
    e, e0 = mg.qp_state()["ebar"], ref.qp_state()["ebar"]
    print(f"hole {hole} {layout}: Newton {len(sm.records)} vs {len(recs0)}, u error {err:.2e}, "
          f"ebar error {np.abs(e - e0).max() / e0.max():.2e}, yielded {(e0 > 0).sum()} of {e0.size}")
    assert err < 1e-9, err
    assert (e0 > 0).sum() > 0.1 * e0.size
    assert np.array_equal(e > 0, e0 > 0)
    assert np.abs(e - e0).max() <= 1e-9 * e0.max()


SPMV_SHAPES = [
    (32, 0), (16, 0), (32, neon_amd.SPMV_NONTEMPORAL), (16, neon_amd.SPMV_NONTEMPORAL),
    (32, neon_amd.SPMV_SLICED), (32, neon_amd.SPMV_SLICED | neon_amd.SPMV_NONTEMPORAL),
    (32, neon_amd.SPMV_SLICED | neon_amd.SPMV_NONTEMPORAL | neon_amd.SPMV_INDEX32),
    (32, neon_amd.SPMV_SLICED | neon_amd.SPMV_INDEX32),
    (32, neon_amd.SPMV_SLICED | neon_amd.SPMV_NONTEMPORAL | neon_amd.SPMV_SYMMETRIC),
    (32, neon_amd.SPMV_SLICED | neon_amd.SPMV_SYMMETRIC),
]


# The sliced copy's two kinds of slice: on the 12^3-cell block (13-node rows) every 64-node
# slice spans several rows and holds more than two node signatures, so all steps keep per-lane
# offsets; on the 120 x 3 x 3 box (121-node rows) two thirds of the slices lie in one row, or
# straddle two rows of one kind, and share each step's offset but for at most two exception
# lanes (the row ends).
SLICED_MESHES = [dict(n=12), dict(n=12, dims=(120, 3, 3))]


def _sliced_mesh(ctx, spec):
    p = problems.block_rollers(**spec)
    mg = neon_amd.DeviceMesh(p.X, p.conn, p.params, context=ctx)
    info = mg.sliced_info()
    assert info["index16"] and info["slices"] == (p.n_nodes + 63) // 64, info
    assert mg.symmetric_info()["usable"]
    if "dims" in spec:
        assert 0.6 * info["steps"] < info["shared_steps"] < info["steps"], info
    else:
        assert info["shared_steps"] == 0, info
    return p, mg


@pytest.mark.parametrize("spec", SLICED_MESHES)
@pytest.mark.parametrize("precision", [neon_amd.PRECISION_FP64, neon_amd.PRECISION_MIXED])
def test_spmv_products_match_the_matrix(ctx, precision, spec):
    # one product y = K x in every shape (and the solvers' instance with x . y partials) against
    # scipy's product with the downloaded values: FP64 to 1e-13 of |K||x| per row, FP32 (x
    # rounded to FP32, FP32 values and accumulation) to 16 FP32 ulps of |K||x| per row, which
    # a misplaced column or value exceeds by orders of magnitude
    import ctypes as C

    from neon_amd import _capi

    p, mg = _sliced_mesh(ctx, spec)
    mg.set_displacement(_smooth_state(p, 0.05, 7))
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
    for lanes, flags in SPMV_SHAPES:
        for extra in (0, _capi.MEASURE_DOT):
            for blocks in (0, 3):
                y = np.zeros(p.n_dofs)
                ctx.check(lib.neon_mesh_spmv_apply(ctx.handle, precision, lanes, flags | extra, blocks,
                                                   _capi.f64(np.ascontiguousarray(x)), _capi.f64(y)))
                err = (np.abs(y - ref) / scale).max()
                print("shape", lanes, flags | extra, blocks, "max error / (|K||x|)", err)
                assert err <= tol, (lanes, flags, extra, blocks, err)


@pytest.mark.parametrize("spec", SLICED_MESHES)
@pytest.mark.parametrize("precision", [neon_amd.PRECISION_FP64, neon_amd.PRECISION_MIXED])
def test_spmv_shapes_agree(ctx, precision, spec):
    # every SpMV shape (CSR-order values with 16 or 32 lanes per block row, the sliced copy with
    # 16- or 32-bit per-lane offsets, plain or nontemporal value loads) solves the device tangent
    # to the same answer as the plain shape and with the same Krylov count up to rounding; each
    # solve's true residual meets the tolerance
    p, mg = _sliced_mesh(ctx, spec)
    u = _smooth_state(p, 0.05, 7)
    mg.set_displacement(u)
    mg.update_internal_variables()
    mg.assemble(precision)
    _, f = mg.download(stiffness=False)
    b = -f.copy()
    mg.enforce_dirichlet(p.fixed_dofs(), b)
    rp, ci = mg.pattern()
    val, _ = mg.download(internal_force=False)
    A = sp.csr_matrix((val, ci, rp))
    results, keys = [], []
    for lanes, flags in SPMV_SHAPES:
        # sliced forms: also fewer workgroups than slices (the grid-stride loops)
        for blocks in (0, 64, 3, 8) if flags & neon_amd.SPMV_SLICED else (0, 64):
            keys.append((lanes, flags, blocks))
            mg.spmv_config(lanes, flags, blocks, blocks)
            x, its, _ = mg.solve(b, criterion=neon_amd.STOP_RELATIVE, tol=1e-10)
            assert np.linalg.norm(b - A @ x) <= 1e-10 * np.linalg.norm(b), (lanes, flags, blocks)
            print("shape", lanes, flags, blocks, "iterations", its)
            results.append((x, its))
    x0, its0 = results[0]
    # the sliced variants on one grid (16/32-bit offsets, plain/nontemporal loads) do the same
    # arithmetic in the same order: equal counts and equal solutions; so do the two symmetric
    # shapes (the symmetric copy's loads do not depend on the nontemporal flag)
    for blocks in (0, 64, 3, 8):
        for symmetric, count in ((0, 4), (neon_amd.SPMV_SYMMETRIC, 2)):
            same = [results[i] for i, key in enumerate(keys)
                    if key[1] & neon_amd.SPMV_SLICED and key[1] & neon_amd.SPMV_SYMMETRIC == symmetric
                    and key[2] == blocks]
            assert len(same) == count
            for x, its in same[1:]:
                assert its == same[0][1] and np.array_equal(x, same[0][0]), blocks
    if precision == neon_amd.PRECISION_FP64:
        slack = 2
    else:
        # the FP32 recurrence of the mixed solver is sensitive to rounding: to the order of the
        # FP32 row sums (the CSR-order kernels sum a row over a lane tree, the sliced one in four
        # interleaved sets along the lane) and of the FP64 dot-product totals.  On the 12^3
        # block the CSR-order shapes take 279 iterations and the sliced copy 303 to 311 as the
        # order of the p.Ap total changed, all with products within one FP32 ulp of |K||x|
        # (test_spmv_products_match_the_matrix) and solutions that agree (below); the counts
        # may differ by 15 %
        slack = max(8, int(0.15 * its0))
    for x, its in results[1:]:
        assert abs(its - its0) <= slack, (its, its0)
        assert np.linalg.norm(x - x0) <= 1e-8 * np.linalg.norm(x0)
    mg.spmv_config()


def test_sliced_columns_without_16bit_offsets(ctx):
    # node numbers shuffled over a 32^3-cell block (35,937 nodes): neighbours lie up to
    # 35,936 apart, beyond 16-bit offsets, so the sliced copy keeps 32-bit per-lane offsets
    # (and no two nodes share a signature, so no step is shared); the solve still meets the
    # tolerance against the downloaded matrix, and agrees with the CSR-order shape.  The
    # unshuffled block uses 16-bit offsets
    p = problems.block_rollers(32)
    perm = np.random.default_rng(3).permutation(p.n_nodes)
    inv = np.argsort(perm)
    X = p.X[perm]
    conn = inv[p.conn.reshape(-1, 8)].astype(np.int32)
    plain = neon_amd.DeviceMesh(p.X, p.conn, p.params, context=ctx)
    assert plain.sliced_info()["index16"]
    mg = neon_amd.DeviceMesh(X, conn, p.params, context=ctx)
    info = mg.sliced_info()
    assert not info["index16"] and info["slices"] == (p.n_nodes + 63) // 64, info
    assert info["shared_steps"] == 0, info
    u = np.zeros(p.n_dofs)
    u.reshape(-1, 3)[:, 2] = 0.01 * X[:, 2]
    mg.set_displacement(u)
    mg.update_internal_variables()
    mg.assemble(neon_amd.PRECISION_FP64)
    _, f = mg.download(stiffness=False)
    b = -f.copy()
    fixed = np.nonzero(np.isin(np.arange(p.n_dofs) // 3, inv[np.nonzero(p.X[:, 2] == 0)[0]]))[0]
    mg.enforce_dirichlet(fixed, b)
    rp, ci = mg.pattern()
    val, _ = mg.download(internal_force=False)
    A = sp.csr_matrix((val, ci, rp))
    x1, its1, _ = mg.solve(b, criterion=neon_amd.STOP_RELATIVE, tol=1e-10)
    assert np.linalg.norm(b - A @ x1) <= 1e-10 * np.linalg.norm(b)
    mg.spmv_config(32, neon_amd.SPMV_NONTEMPORAL, 0, 0)
    x2, its2, _ = mg.solve(b, criterion=neon_amd.STOP_RELATIVE, tol=1e-10)
    assert abs(its1 - its2) <= 2 and np.linalg.norm(x1 - x2) <= 1e-8 * np.linalg.norm(x2)
