"""The HIP path at the workloads' full sizes (64^3 FP64 block, 96^3 adjoint block) and the J2
notched plate at half size (64 x 32 x 32 cells, 20 load steps), through properties that need
no CPU oracle run (see full_size_checks.py for what each checks and why)."""
import pytest

import full_size_checks as fsc

pytestmark = pytest.mark.gpu


def test_fp64_block_full_size(ctx):
    r = fsc.fp64_block(ctx, 64)
    print(r)
    assert r["fp64_status_ok"] and r["mixed_status_ok"]
    assert len(r["fp64_newton_per_step"]) == 10
    # same Newton count per load step in both precisions
    assert r["mixed_newton_per_step"] == r["fp64_newton_per_step"]
    assert r["mixed_vs_fp64"] < 1e-8
    # converged: the residual on the free dofs is small against the reactions (Newton's
    # relative tolerance is 1e-8)
    assert r["free_residual"] < 1e-8
    # every element's forces balance (shape functions sum to one), to roundoff
    assert r["force_balance"] < 1e-10


def test_adjoint_block_full_size(ctx):
    r = fsc.adjoint_block(ctx, 96, directions=2)
    print(r)
    assert max(r["fd_relative_errors"]) < 1e-6


def test_j2_plate_half_size(ctx):
    r = fsc.j2_plate(ctx, 64, 20, rel_steps=(1e-7,))
    print(r)
    assert r["yielded_qps"] > 0
    assert r["fd_relative_error"][1e-7] < 1e-6
