"""Roofline probe of the node-block SpMV on the assembled tangent of the neo-Hooke block.

    python -m neon_amd.roofline --n 128 --precision mixed --reps 20

builds the n^3 block (the bench's mesh), assembles the tangent at u = 0 in the requested
precision, runs `reps` SpMVs in the solvers' form (`neon_mesh_spmv_bench` of the measurement library
libneon_hip_measure.so: the sliced copy, nontemporal value loads, x . y block partials; the CG
loop's kernel instance; the symmetric copy where the mesh allows it unless --full) and a
streaming copy of `--copy-mb` MiB (`neon_stream_probe`, the
calibration kernel for the FETCH_SIZE / WRITE_SIZE counters), and prints one JSON line: the
mean SpMV time, its algorithmic bytes (`spmv_bytes` or `symmetric_bytes`) and the achieved
fraction of the 8 TB/s HBM peak.  With --assembly it times the tangent + f_int assembly instead
(its canonical flops and algorithmic bytes).  Run it under `rocprofv3 --kernel-trace
--stats` or `--pmc` to read the same kernels' durations and traffic counters.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json

import numpy as np

import neon_amd
from neon_amd import _capi, problems

HBM_PEAK_GBS = 8000.0


def spmv_bytes(nnz: int, nodes: int, value_bytes: int) -> int:
    """Algorithmic bytes of one node-block SpMV: the nonzero values once, one 16-bit column
    offset per 3x3 block, x read once, y written once (as bench.py counts them)."""
    n = 3 * nodes
    return value_bytes * nnz + 2 * (nnz // 9) + 2 * value_bytes * n


def symmetric_bytes(nnz: int, nodes: int, value_bytes: int) -> int:
    """Algorithmic bytes of one product on the symmetric copy: the diagonal and upper blocks'
    values once ((nnz + 9 nodes) / 2), x read once, y written once (as bench.py counts them)."""
    n = 3 * nodes
    return value_bytes * (nnz + 9 * nodes) // 2 + 2 * value_bytes * n


ASSEMBLY_FLOP_PER_ELEMENT = 45536  # 8 QPs x (B^T D B 4320 + geometric 528 + kinematics 576 + f_int 168 + constitutive 100)
FP64_PEAK_TFLOPS = 78.6


def assembly_bytes(nnz: int, nodes: int, elems: int, value_bytes: int) -> int:
    """Algorithmic bytes of one assembly: the tangent written once, coordinates and displacement
    (48 per node), connectivity and parameters (40 per element), f_int (8 per dof)."""
    return value_bytes * nnz + 48 * nodes + 40 * elems + 8 * 3 * nodes


def assembly(args, ctx, lib, p, mesh, prec) -> dict:
    """One assembly (all colours, tangent + f_int) of the held state, timed over `reps`, with the
    streaming copy that calibrates FETCH_SIZE / WRITE_SIZE in the same process."""
    ms = C.c_double()
    ctx.check(lib.neon_mesh_assembly_bench(ctx.handle, prec, 128, args.reps, C.byref(ms)))
    vb = 8 if prec == _capi.PRECISION_FP64 else 4
    nnz = mesh.nnz()
    b = assembly_bytes(nnz, p.n_nodes, p.n_elem, vb)
    flop = ASSEMBLY_FLOP_PER_ELEMENT * p.n_elem
    read_ms, copy_ms = C.c_double(), C.c_double()
    copy_bytes = args.copy_mb << 20
    ctx.check(lib.neon_stream_probe(ctx.handle, copy_bytes, args.reps, C.byref(read_ms), C.byref(copy_ms)))
    out = {
        "workload": f"neo-Hooke block {args.n}^3, assembly at u = 0",
        "precision": args.precision,
        "elements": p.n_elem,
        "nnz": nnz,
        "assembly_avg_ms": ms.value,
        "assembly_flop": flop,
        "assembly_tflops": flop / (ms.value * 1e-3) / 1e12,
        "fp64_peak_frac": flop / (ms.value * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
        "assembly_algorithmic_bytes": b,
        "copy_bytes_each_way": copy_bytes,
    }
    print(json.dumps(out), flush=True)
    return out


def main(argv=None) -> dict:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--copy-mb", type=int, default=1024)
    ap.add_argument("--dirty", action="store_true")
    ap.add_argument("--full", action="store_true", help="the full sliced copy even where the symmetric one applies")
    ap.add_argument("--assembly", action="store_true",
                    help="time the tangent + f_int assembly (neon_mesh_assembly_bench) instead of the SpMV")
    args = ap.parse_args(argv)

    ctx = neon_amd.Context(0)
    p = problems.block_rollers(args.n)
    mesh = neon_amd.DeviceMesh(p.X, p.conn, p.params, context=ctx)
    mesh.set_displacement(np.zeros(p.n_dofs))
    mesh.update_internal_variables()
    prec = _capi.PRECISION_FP64 if args.precision == "fp64" else _capi.PRECISION_MIXED
    mesh.assemble(prec)
    ms = C.c_double()
    lib = _capi.measure_lib()
    if args.assembly:
        return assembly(args, ctx, lib, p, mesh, prec)
    # the solvers' form exactly: the sliced copy (symmetric where usable) with nontemporal value
    # loads and the x . y block partials (the same kernel instance as inside the CG loop),
    # workgroups by size; --dirty puts a write of 4 vectors of n before every product, as the CG
    # vector updates do (timed apart)
    symmetric = mesh.symmetric_info()
    sym = symmetric["usable"] and not args.full
    flags = _capi.SPMV_SLICED | _capi.SPMV_NONTEMPORAL | _capi.MEASURE_DOT
    if sym:
        flags |= _capi.SPMV_SYMMETRIC
    if args.dirty:
        flags |= _capi.MEASURE_DIRTY
    cfg = dict(lanes=32, flags=flags, blocks=0)
    ctx.check(lib.neon_mesh_spmv_bench(ctx.handle, prec, cfg["lanes"], cfg["flags"], cfg["blocks"], args.reps,
                                       C.byref(ms)))
    nnz = mesh.nnz()
    vb = 8 if prec == _capi.PRECISION_FP64 else 4
    sliced = mesh.sliced_info()
    b = symmetric_bytes(nnz, p.n_nodes, vb) if sym else spmv_bytes(nnz, p.n_nodes, vb)
    gbs = b / (ms.value * 1e-3) / 1e9
    read_ms, copy_ms = C.c_double(), C.c_double()
    copy_bytes = args.copy_mb << 20
    ctx.check(lib.neon_stream_probe(ctx.handle, copy_bytes, args.reps, C.byref(read_ms), C.byref(copy_ms)))
    out = {
        "workload": f"neo-Hooke block {args.n}^3, tangent at u = 0",
        "precision": args.precision,
        "nodes": p.n_nodes,
        "dofs": p.n_dofs,
        "nnz": nnz,
        "shape": cfg,
        "sliced": sliced,
        "symmetric": dict(symmetric, used=sym),
        "spmv_avg_ms": ms.value,
        "spmv_algorithmic_bytes": b,
        "spmv_gbs": gbs,
        "spmv_hbm_frac": gbs / HBM_PEAK_GBS,
        "copy_bytes_each_way": copy_bytes,
        "stream_read_gbs": copy_bytes / (read_ms.value * 1e-3) / 1e9,
        "stream_copy_gbs": 2 * copy_bytes / (copy_ms.value * 1e-3) / 1e9,
    }
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
