"""Checks of the HIP path at the workloads' full sizes, through properties that do not need
the CPU oracle (which would take minutes to hours there).

* FP64 block (64^3 neo-Hooke, jittered, per-element mu, 10 load steps to 20 % stretch): the
  FP64 and mixed-precision Newton runs converge every step with the same Newton counts and
  agree on u; at the converged state the internal force on the free dofs (the residual) is
  small against the reactions on the constrained ones, and its total over all dofs vanishes
  (the shape functions sum to one, so every element's forces balance; a check of assembly
  at any state).
* Mixed-precision block (128^3, 6.44 M dofs): load step 1 to the 1e-8 relative residual in
  FP64 and mixed precision; the same Newton count, both residuals at the target, and the
  mixed displacement against the FP64 one.
* Adjoint block (96^3, smoothed log-normal mu, 4096 observed top nodes): the adjoint gradient
  in log mu against FP64 central differences (step 1e-6) along seeded random directions.
* J2 notched plate (128 x 64 x 64 cells less the hole of radius 8h, 20 load steps to 0.5 %
  strain): the checkpointed adjoint gradient in the yield stresses against central
  differences of two full forward runs, along the normalised gradient (a random direction's
  derivative is ~sqrt(elements) times smaller and drowns in the forward solves' roundoff at
  this size).  The plastic response has kinks where quadrature points cross the yield
  surface; a relative step h lets ~h of them cross inside the difference, so the error
  falls with h (on the half-size plate 1.0e-6, 3.3e-7, 6.0e-8, 5.7e-9 at h = 1e-6, 3e-7,
  1e-7, 3e-8, with no roundoff floor yet), and the checks use h = 1e-7.

Run as `python tests/full_size_checks.py [fp64_block mixed_block_step1 adjoint_block j2_plate
adjoint_block32]` for a JSON line per check (adjoint_block32: the 32 directions of the
adjoint workload); `tests/test_gpu_full_size.py` runs each at its workload's size (the adjoint
block along 8 directions).
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import neon_amd  # noqa: E402
from neon_amd import problems  # noqa: E402

import adjoint_harness as ah  # noqa: E402


def _progress(*a):
    # long runs print as they go (a forward run of the full plate takes about a minute)
    print(*a, file=sys.stderr, flush=True)


def _per_step_newton(records):
    out = {}
    for r in records:
        key = round(r["step_time"], 12)
        out[key] = out.get(key, 0) + 1
    return out


def fp64_block(ctx, n=64):
    """FP64 and mixed Newton runs of the roller block; returns a summary dict."""
    p = problems.block_rollers(n)
    res = {"check": "fp64_block", "n": n, "dofs": p.n_dofs}
    sols = {}
    for prec in ("fp64", "mixed"):
        mesh = neon_amd.DeviceMesh(p.X, p.conn, p.params, context=ctx)
        sim = dict(p.simulation)
        sim["linear_solver"] = dict(sim["linear_solver"], backend={"type": "hip", "precision": prec})
        sm = neon_amd.static_matrix(mesh, sim, p.boundaries)
        t = time.time()
        u = sm.solve()
        res[f"{prec}_seconds"] = time.time() - t
        res[f"{prec}_newton_per_step"] = _per_step_newton(sm.records)
        res[f"{prec}_status_ok"] = all(r["status"] == 0 for r in sm.records)
        sols[prec] = (u, mesh)
    u64, mesh64 = sols["fp64"]
    res["mixed_vs_fp64"] = float(np.abs(sols["mixed"][0] - u64).max() / np.abs(u64).max())
    # internal force at the converged FP64 state: residual on the free dofs against the
    # reactions; total over all dofs against the pulled (top) face's reaction
    mesh64.set_displacement(u64)
    mesh64.update_internal_variables()
    mesh64.assemble(neon_amd.PRECISION_FP64)
    _, f = mesh64.download(stiffness=False)
    top = np.asarray(p.boundaries[-1]["dofs"])
    fixed = np.zeros(p.n_dofs, dtype=bool)
    fixed[p.fixed_dofs()] = True
    res["top_reaction"] = float(f[top].sum())
    res["free_residual"] = float(np.linalg.norm(f[~fixed]) / np.linalg.norm(f[fixed]))
    res["force_balance"] = float(abs(f[2::3].sum()) / abs(res["top_reaction"]))
    return res


def mixed_block_step1(ctx, n=128):
    """The mixed-precision workload (128^3 roller block, 6.44 M dofs): load step 1 to the 1e-8
    relative Newton residual in FP64 and in mixed precision on one mesh; Newton counts, force
    norms and the mixed path's displacement error against the FP64 path."""
    p = problems.block_rollers(n, residual_tolerance=1e-8, displacement_tolerance=1e-8, norm_mode=1,
                               solver_tolerance=1e-10, max_iterations=200000)
    mesh = neon_amd.DeviceMesh(p.X, p.conn, p.params, context=ctx)
    res = {"check": "mixed_block_step1", "n": n, "dofs": p.n_dofs}
    sols = {}
    for prec in ("fp64", "mixed"):
        sim = dict(p.simulation, max_load_steps=1)
        sim["linear_solver"] = dict(sim["linear_solver"], backend={"type": "hip", "precision": prec})
        sm = neon_amd.static_matrix(mesh, sim, p.boundaries)
        t = time.time()
        sols[prec] = sm.solve()
        res[f"{prec}_seconds"] = time.time() - t
        res[f"{prec}_newton_iterations"] = len(sm.records)
        res[f"{prec}_force_norms"] = [r["force_norm"] for r in sm.records]
        res[f"{prec}_displacement_norms"] = [r["displacement_norm"] for r in sm.records]
        res[f"{prec}_cg_iterations"] = [r["solver_iterations"] for r in sm.records]
        res[f"{prec}_status_ok"] = all(r["status"] == 0 for r in sm.records)
        _progress(f"mixed_block_step1 {n}: {prec} {res[f'{prec}_seconds']:.1f} s, {len(sm.records)} Newton iterations")
    u64, u32 = sols["fp64"], sols["mixed"]
    res["mixed_vs_fp64_max"] = float(np.abs(u32 - u64).max() / np.abs(u64).max())
    res["mixed_vs_fp64_l2"] = float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))
    return res


def adjoint_block(ctx, n=96, directions=2):
    """Adjoint gradient against central differences on the adjoint block."""
    p = problems.adjoint_block(n, stretch=0.05, n_obs=4096)
    log_mu = np.log(p.extra["mu"])
    t0 = time.time()
    _, target = ah.gpu_forward(p, np.log(p.extra["mu_target"]), ctx)
    mesh, u = ah.gpu_forward(p, log_mu, ctx)
    dofs = ah.observed_dofs(p)
    t = time.time()
    grad4 = mesh.adjoint_gradient(ah.objective_gradient(u, dofs, target), p.fixed_dofs(), tol=1e-13)
    res = {"check": "adjoint_block", "n": n, "elements": p.n_elem, "dofs": p.n_dofs,
           "adjoint_seconds": time.time() - t, "adjoint_cg_iterations": mesh.adjoint_iterations}
    g = ah.log_mu_gradient(grad4, log_mu)
    rel = []
    for k, d in enumerate(ah.directions(p.n_elem, directions)):
        fd = ah.gpu_fd(p, log_mu, u, d, target, ctx)
        rel.append(abs(fd - g @ d) / abs(g @ d))
        _progress(f"adjoint_block {n}: direction {k + 1} of {directions}: {rel[-1]:.2e} at {time.time() - t0:.1f} s")
    res["fd_relative_errors"] = rel
    res["seconds"] = time.time() - t0
    return res


def j2_plate(ctx, nx=128, steps=20, rel_steps=(1e-7,)):
    """Checkpointed J2 adjoint against central differences along the normalised gradient,
    one pair of forward runs per relative step in `rel_steps`."""
    ny = nx // 2
    p = problems.notched_plate_j2(nx, ny, ny, steps=steps, hole_cells=8.0 * ny / 64, residual_tolerance=1e-10,
                                  displacement_tolerance=1e-10, norm_mode=1, solver_tolerance=1e-13,
                                  max_iterations=100000)
    p.simulation["record_trajectory"] = True
    w = np.random.default_rng(21).normal(size=p.n_dofs)
    w[p.fixed_dofs()] = 0.0

    def forward(params):
        mesh = neon_amd.DeviceMesh(p.X, p.conn, params, model=neon_amd.MODEL_J2,
                                   tangent=neon_amd.TANGENT_CONSISTENT, context=ctx)
        sm = neon_amd.static_matrix(mesh, p.simulation, p.boundaries)
        return mesh, sm.solve(), sm.records

    t0 = time.time()
    mesh, u, recs = forward(p.params)
    _progress(f"j2_plate {nx}: forward {time.time() - t0:.1f} s")
    res = {"check": "j2_plate", "cells": [nx, ny, ny], "elements": p.n_elem, "dofs": p.n_dofs,
           "forward_seconds": time.time() - t0, "newton_iterations": len(recs),
           "cg_iterations": int(sum(r["solver_iterations"] for r in recs)),
           "yielded_qps": int((mesh.qp_state()["ebar"] > 0).sum())}
    t = time.time()
    g = mesh.adjoint_gradient(w, p.fixed_dofs(), tol=1e-14)
    res["adjoint_seconds"] = time.time() - t
    res["adjoint_cg_iterations"] = mesh.adjoint_iterations
    _progress(f"j2_plate {nx}: adjoint {res['adjoint_seconds']:.1f} s")
    gy = g[:, 2]
    d = gy / np.sqrt(np.mean(gy * gy))  # unit RMS entries along the gradient
    res["adjoint"] = float(gy @ d)
    res["fd"], res["fd_relative_error"] = {}, {}
    for rel in rel_steps:
        h = rel * np.abs(p.params[:, 2]).mean()
        vals = []
        for sgn in (1.0, -1.0):
            pp = p.params.copy()
            pp[:, 2] += sgn * h * d
            vals.append(float(w @ forward(pp)[1]))
            _progress(f"j2_plate {nx}: step {rel:g} side {sgn:+.0f} done at {time.time() - t0:.1f} s")
        fd = (vals[0] - vals[1]) / (2.0 * h)
        res["fd"][rel] = fd
        res["fd_relative_error"][rel] = abs(fd - gy @ d) / abs(gy @ d)
    res["seconds"] = time.time() - t0
    return res


def main(argv):
    names = argv or ["fp64_block", "mixed_block_step1", "adjoint_block", "j2_plate"]
    ctx = neon_amd.Context(0)
    for name in names:
        if name == "adjoint_block32":
            r = adjoint_block(ctx, 96, directions=32)
        else:
            r = {"fp64_block": fp64_block, "mixed_block_step1": mixed_block_step1, "adjoint_block": adjoint_block,
                 "j2_plate": j2_plate}[name](ctx)
        print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
