"""Benchmark: implicit Newton steps/s on the 128^3 jittered neo-Hooke block (6.44 M dofs),
mixed precision (headline) and FP64, each Newton step = assembly + Dirichlet + Jacobi-PCG to
1e-10 relative + state update, relative Newton residual target 1e-8.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--no-cpu-baseline] [--layout full]

Prints one JSON line.  `value` is the mixed-precision Newton steps/s over exactly K timed
Newton iterations (inputs resident in HBM); `roofline` is the FP32 node-block SpMV of the
mixed inner solve, timed live with HIP events on the solver's stream; `cpu_baseline` is the
CPU FP64 oracle timed on a bounded sample on this host.
"""
from __future__ import annotations

import argparse
import json
import subprocess
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
ASSEMBLY_FLOP_PER_ELEMENT = 45536  # canonical count: 8 QPs x (B^T D B 4320 + geometric 528 +
#                                    kinematics 576 + f_int 168 + constitutive 100)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_PROFILES = {False: os.path.join(ROOT, "profiles", "spmv_sell_neohooke128_traffic.txt"),
                    True: os.path.join(ROOT, "profiles", "spmv_sym_neohooke128_traffic.txt")}


def measured_traffic(n, symmetric):
    """HBM bytes per launch of the FP32 node-block SpMV (the symmetric or the full sliced copy)
    at the 128^3 workload, from the calibrated rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    summarised in the profile file (None for other sizes or without the file)."""
    path = TRAFFIC_PROFILES[bool(symmetric)]
    if n != 128 or not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith("traffic_bytes_per_launch_fp32"):
                return int(line.split("=")[1])
    return None


def spmv_bytes(nnz, n, value_bytes):
    """Algorithmic bytes of one node-block SpMV: the nonzero values once, one 16-bit column
    offset per 3x3 block (nnz / 9 of them), x read once, y written once.  The sliced copy the
    solvers read stores most column offsets once per step of 64 nodes instead of once per
    block, and pads short nodes with zero values; neither is counted (`sliced_bytes` gives
    what the copy holds)."""
    return value_bytes * nnz + 2 * (nnz // 9) + 2 * value_bytes * n


def sliced_bytes(n, value_bytes, sliced):
    """Bytes one product on the sliced copy reads and writes (`sliced` = ``DeviceMesh.sliced_info()``):
    the step records (9 x 64 values per step, padding included), a code and an exception word
    per step, 64 per-lane offsets (2 or 4 bytes) per step without a shared offset, the
    lane-to-node map, the exception lanes and step pointer of each slice, x once and y once."""
    ib = 2 if sliced["index16"] else 4
    own = sliced["steps"] - sliced["shared_steps"]
    meta = 8 * sliced["steps"] + 64 * ib * own + (4 * 64 + 8) * sliced["slices"] + 4
    return value_bytes * 576 * sliced["steps"] + meta + 2 * value_bytes * n


def symmetric_bytes(nnz, n, value_bytes):
    """Algorithmic bytes of one product on the symmetric sliced copy: the stored values once
    (the diagonal block and the upper blocks of each node, (nnz + 3 n) / 2 values), x read once,
    y written once.  The lower blocks are the stored upper ones read a second time, from the L2
    where the schedule keeps them there: not counted (the measured traffic shows what was)."""
    return value_bytes * (nnz + 3 * n) // 2 + 2 * value_bytes * n


def symmetric_copy_bytes(n, value_bytes, symmetric, slices):
    """Bytes the symmetric copy holds and one product touches once: 1 + offsets records of
    9 x 64 values per slice (padding included), the schedule's wave slots, x and y."""
    return value_bytes * 576 * slices * (1 + symmetric["offsets"]) + 4 * symmetric["slots"] + 2 * value_bytes * n


def standalone_products(ctx, mesh, nnz, n, reps=20):
    """The FP32 products on the mesh's assembled mixed-precision tangent, outside the timed
    region, through the measurement library (solver form: p.Ap partials, each product after a
    write of 4 vectors): the symmetric copy (the solvers' choice where it applies) and the full
    sliced copy, so that the two layouts' times and roofline fractions come from the same run."""
    import ctypes as C

    from neon_amd import _capi

    lib = _capi.measure_lib()
    base = _capi.SPMV_SLICED | _capi.SPMV_NONTEMPORAL | _capi.MEASURE_DOT | _capi.MEASURE_DIRTY
    out = {}
    for name, flags, nbytes in (("symmetric_copy", base | _capi.SPMV_SYMMETRIC, symmetric_bytes(nnz, n, 4)),
                                ("full_copy", base, spmv_bytes(nnz, n, 4))):
        ms = C.c_double()
        ctx.check(lib.neon_mesh_spmv_bench(ctx.handle, _capi.PRECISION_MIXED, 32, flags, 0, reps, C.byref(ms)))
        gbs = nbytes / (ms.value * 1e-3) / 1e9
        out[name] = {"avg_ms": ms.value, "algorithmic_bytes": nbytes, "achieved": gbs, "frac": gbs / HBM_PEAK_GBS}
    return out


def run_precision(neon_amd, problems, ctx, mesh, p, precision, steps, warmup):
    sim = dict(p.simulation)
    sim["linear_solver"] = dict(sim["linear_solver"], backend={"type": "hip", "precision": precision})
    from neon_amd import _capi

    if warmup > 0:
        sm = neon_amd.static_matrix(mesh, dict(sim, max_total_iterations=warmup), p.boundaries)
        sm.solve()
    spmv_class = _capi.KERNEL_SPMV_FP32 if precision == "mixed" else _capi.KERNEL_SPMV_FP64
    ctx.profile(True)
    for c in range(4):
        ctx.profile_read(c, reset=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    sm = neon_amd.static_matrix(mesh, dict(sim, max_total_iterations=steps), p.boundaries)
    sm.solve()
    ctx.synchronize()
    t1 = time.perf_counter()
    spmv_count, spmv_ms = ctx.profile_read(spmv_class)
    asm_count, asm_ms = ctx.profile_read(_capi.KERNEL_ASSEMBLY)
    ctx.profile(False)
    timed = dict(seconds=t1 - t0, records=sm.records, timing=sm.timing, spmv_count=spmv_count, spmv_ms=spmv_ms,
                 asm_count=asm_count, asm_ms=asm_ms,
                 load_steps=sorted({r["step"] + 1 for r in sm.records}))
    log(f"[{precision}] {steps} Newton steps in {t1 - t0:.3f} s; CG its "
        f"{[r['solver_iterations'] for r in sm.records]}; spmv {spmv_count} x {spmv_ms / max(spmv_count, 1):.4f} ms")
    # first load step to the 1e-8 relative residual
    t0 = time.perf_counter()
    sm = neon_amd.static_matrix(mesh, dict(sim, max_load_steps=1), p.boundaries)
    sm.solve()
    t1 = time.perf_counter()
    u1 = mesh.displacement()
    timed["to_tolerance"] = dict(seconds=t1 - t0, newton_iterations=len(sm.records),
                                 final_force_norm=sm.records[-1]["force_norm"],
                                 final_displacement_norm=sm.records[-1]["displacement_norm"],
                                 cg_iterations=[r["solver_iterations"] for r in sm.records],
                                 reliable_updates=sm.timing.get("refinement_count", 0))
    timed["u_step1"] = u1
    log(f"[{precision}] load step 1 to 1e-8: {len(sm.records)} Newton its, {t1 - t0:.3f} s, "
        f"force norms {['%.2e' % r['force_norm'] for r in sm.records]}")
    return timed


def cpu_baseline(problems, n_gpu, cg_its_per_step, nnz_gpu, n_elem_gpu):
    """Oracle (C++ FP64 port of the reference CPU path) on a 64^3 sample of the same
    recipe: three state updates + assemblies (parallel element loop, binary-search + atomic
    scatter) + serial f_int, and 1500 Jacobi-PCG iterations with an OpenMP SpMV (about 10 s of
    CPU work); scaled to the full mesh by element and nnz counts and to a Newton step by the
    GPU FP64 run's mean CG iterations per Newton step."""
    from oracle import oracle as orc

    threads = cpu_threads()
    ns = 64
    p = problems.block_rollers(ns)
    m = orc.Model(p.X, p.conn, p.params, model=p.model)
    rng = np.random.default_rng(0)
    u = 1e-3 * rng.normal(size=p.n_dofs)
    rp, ci = m.pattern()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        m.update(u)
    t1 = time.perf_counter()
    for _ in range(reps):
        val, f = m.assemble(rp, ci, threads=threads)
    t2 = time.perf_counter()
    b = -f
    orc.enforce_dirichlet(rp, ci, val, b, p.fixed_dofs())
    t3 = time.perf_counter()
    its_sample = 1500
    _, _, its, _ = orc.cg_eigen(rp, ci, val, b, 1e-300, its_sample, threads=threads)
    t4 = time.perf_counter()
    per_it = (t4 - t3) / max(its, 1)
    scale_e = n_elem_gpu / p.n_elem
    scale_nnz = nnz_gpu / len(ci)
    step = (t2 - t0) / reps * scale_e + (t3 - t2) * scale_nnz + per_it * scale_nnz * cg_its_per_step
    spmv_b = 12 * len(ci) + 4 * (p.n_dofs + 1) + 16 * p.n_dofs
    return {
        "value": 1.0 / step,
        "unit": "Newton steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{ns}^3 block ({p.n_elem} elements, {len(ci)} nnz): {reps} state updates + assemblies + f_int "
                   f"({t2 - t0:.2f} s) and {its} Jacobi-PCG iterations ({t4 - t3:.2f} s), extrapolated to "
                   f"{n_elem_gpu} elements / {nnz_gpu} nnz and {cg_its_per_step:.0f} CG iterations per Newton step"),
        "cg_iteration_ms_full_mesh": 1e3 * per_it * scale_nnz,
        "spmv_gbs": spmv_b / per_it / 1e9,
        "host": host_info(),
    }


def _span(steps):
    return f"{steps[0]}-{steps[-1]}" if len(steps) > 1 else str(steps[0]) if steps else "none"


def cpu_threads():
    """Every core of this process's affinity set, bounded by the job's CPU share when one is
    set (OMP_NUM_THREADS: the host grants a one-GPU job that many cores)."""
    cores = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        cores = min(cores, int(share))
    return cores


def host_info():
    """nproc, the affinity set, the job's CPU share and the lscpu model name of this host."""
    model = "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    cpuinfo = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpuinfo = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "threads_used": cpu_threads(),
            "lscpu_model": model, "cpuinfo_model": cpuinfo}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precisions", default="fp64,mixed")
    ap.add_argument("--layout", choices=["default", "full"], default="default",
                    help="full: the full sliced copy of the tangent even where the symmetric copy applies")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "neon_amd runs on one GPU; the multi-GPU decomposition (z-slabs of elements, "
                                   "one halo exchange and two scalar all-reduces per CG iteration) is a design only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import neon_amd
    from neon_amd import problems

    t0 = time.perf_counter()
    p = problems.block_rollers(args.n, residual_tolerance=1e-8, displacement_tolerance=1e-8, norm_mode=1,
                               solver_tolerance=1e-10, max_iterations=200000)
    ctx = neon_amd.Context(0)
    mesh = neon_amd.DeviceMesh(p.X, p.conn, p.params, model=p.model, context=ctx)
    nnz = mesh.nnz()
    sliced = mesh.sliced_info()
    symmetric = mesh.symmetric_info()
    sym = symmetric["usable"] and args.layout == "default"
    if args.layout == "full":
        mesh.spmv_config(32, neon_amd.SPMV_SLICED | neon_amd.SPMV_NONTEMPORAL)
    log(f"setup {time.perf_counter() - t0:.1f} s: {p.n_elem} elements, {p.n_dofs} dofs, nnz {nnz}")

    results = {}
    for prec in args.precisions.split(","):
        results[prec] = run_precision(neon_amd, problems, ctx, mesh, p, prec, args.steps, args.warmup)

    head = "mixed" if "mixed" in results else "fp64"
    r = results[head]
    value = args.steps / r["seconds"]
    vb = 4 if head == "mixed" else 8

    def layout_bytes(value_bytes):
        return symmetric_bytes(nnz, p.n_dofs, value_bytes) if sym else spmv_bytes(nnz, p.n_dofs, value_bytes)

    bytes_spmv = layout_bytes(vb)
    vname = "float" if vb == 4 else "double"
    if sym:
        kernel = f"spmv_sym_kernel<{vname}> (symmetric sliced node-block SpMV + p.Ap)"
        layout = dict(symmetric, slices=sliced["slices"],
                      bytes_per_launch=symmetric_copy_bytes(p.n_dofs, vb, symmetric, sliced["slices"]))
    else:
        kernel = f"spmv_sell_kernel<{vname}> (sliced node-block SpMV + p.Ap)"
        layout = dict(sliced, bytes_per_launch=sliced_bytes(p.n_dofs, vb, sliced))
    spmv_avg_ms = r["spmv_ms"] / max(r["spmv_count"], 1)
    achieved = bytes_spmv / (spmv_avg_ms * 1e-3) / 1e9
    line = {
        "metric": "Newton steps/s & SpMV GB/s at 6.3M dofs: mixed vs FP64 to 1e-8 residual",
        "value": value,
        "unit": "Newton steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner PCG + f64 refinement/residuals" if head == "mixed" else "f64",
        "data": "synthetic (seeded jittered structured hex8 mesh, per-element mu ~ U[0.5,1.5] MPa)",
        "config": {"workload": f"neo-Hooke block {args.n}^3 hex8, rollers x=0/y=0/z=0, top u_z = 20% in 10 "
                               f"steps; the timed Newton iterations are the first {args.steps} of the run from u = 0 "
                               f"(after {args.warmup} untimed), covering load steps {_span(r['load_steps'])}; "
                               f"Jacobi-PCG 1e-10; Newton target 1e-8",
                   "elements": p.n_elem, "dofs": p.n_dofs, "nnz": nnz, "load_steps_timed": r["load_steps"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.n, sym) if head == "mixed" else None,
                     "kernel": kernel,
                     "algorithmic_bytes": bytes_spmv, "avg_ms": spmv_avg_ms, "launches": r["spmv_count"],
                     "symmetric_layout" if sym else "sliced_layout": layout},
    }
    for prec, rr in results.items():
        vbp = 4 if prec == "mixed" else 8
        bs = layout_bytes(vbp)
        avg = rr["spmv_ms"] / max(rr["spmv_count"], 1)
        asm_avg = rr["asm_ms"] / max(rr["asm_count"], 1)
        its = [x["solver_iterations"] for x in rr["records"]]
        line[prec] = {
            "newton_steps_per_s": args.steps / rr["seconds"],
            "ms_per_newton_step": 1e3 * rr["seconds"] / args.steps,
            "spmv_gbs": bs / (avg * 1e-3) / 1e9,
            "spmv_avg_ms": avg,
            "spmv_hbm_frac": bs / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "cg_iterations_per_newton_step": its,
            "assembly_avg_ms": asm_avg,
            "assembly_gflops": ASSEMBLY_FLOP_PER_ELEMENT * p.n_elem / (asm_avg * 1e-3) / 1e9 if asm_avg > 0 else None,
            "assembly_fp64_peak_frac": (ASSEMBLY_FLOP_PER_ELEMENT * p.n_elem / (asm_avg * 1e-3) / 1e12
                                        / FP64_PEAK_TFLOPS) if asm_avg > 0 else None,
            "phase_ms": {k: rr["timing"][k] for k in ("assemble_ms", "dirichlet_ms", "solve_ms", "update_ms",
                                                      "total_ms")},
            "to_1e-8": rr["to_tolerance"],
            "reliable_updates": rr["timing"].get("refinement_count", 0),
        }
    if "mixed" in results and "fp64" in results:
        line["mixed_vs_fp64_speedup"] = {
            "newton_steps_per_s": line["mixed"]["newton_steps_per_s"] / line["fp64"]["newton_steps_per_s"],
            "time_to_1e-8": results["fp64"]["to_tolerance"]["seconds"] / results["mixed"]["to_tolerance"]["seconds"],
        }
        # the mixed path's error against the FP64 path after load step 1 (both to the 1e-8 residual)
        u64, u32 = results["fp64"]["u_step1"], results["mixed"]["u_step1"]
        line["mixed_vs_fp64_displacement_error"] = {
            "max_abs_rel": float(np.abs(u32 - u64).max() / np.abs(u64).max()),
            "l2_rel": float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64)),
            "load_step": 1,
        }
    if sym and "mixed" in results:
        # both FP32 layouts on the tangent the mixed run left assembled, outside the timed region
        line["roofline"]["standalone_fp32_products"] = standalone_products(ctx, mesh, nnz, p.n_dofs)
    if not args.no_cpu_baseline:
        its64 = [x["solver_iterations"] for x in results.get("fp64", r)["records"]]
        line["cpu_baseline"] = cpu_baseline(problems, args.n, float(np.mean(its64)), nnz, p.n_elem)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
