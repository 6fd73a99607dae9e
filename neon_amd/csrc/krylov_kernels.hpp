// Device kernels of the Jacobi-preconditioned Krylov solvers: CSR SpMV in two layouts,
// fused vector updates with grid reductions, and the device-resident scalar state.
//
// The algorithms follow neon's two solver families:
//   * Eigen's ConjugateGradient behind conjugate_gradient::solve
//     (src/solver/linear/linear_solver.cpp:67-109): relative stop ||r||^2 < tol^2 ||b||^2,
//   * conjugate_gradient_cuda::solve (src/solver/linear/conjugate_gradient_cuda.cpp:95-186):
//     absolute stop sqrt(r.z) <= tol,
// with every scalar kept on the device (the CUDA code synchronises on each cuBLAS dot).
#pragma once

#include "common.hpp"
#include "neon_hip.h"

#include <cfloat>
#include <type_traits>
#include <utility>

namespace neon_amd
{
// Scalars of one Krylov iteration, ping-ponged between two device slots so that the
// kernel that computes iteration k+1's scalars never overwrites what others still read.
struct krylov_state
{
    double rz;        // r . z  (Eigen's absNew, CUDA's "residual")
    double rr;        // r . r
    double rhs2;      // b . b
    double threshold; // Eigen: max(tol^2 ||b||^2, DBL_MIN)
    double alpha, beta, omega, rho; // BiCGSTAB scalars (alpha reused by CG for reports)
    double aux0, aux1;               // reliable-update bookkeeping (mixed precision)
    int32_t it;   // iterations as the reference counts them
    int32_t done; // 0 running, 1 converged, 2 iteration limit, 3 non-finite, 4 breakdown
    int32_t restarts;
    int32_t refine; // mixed precision: a reliable update is due this iteration
};

enum : int32_t
{
    KRYLOV_RUNNING = 0,
    KRYLOV_CONVERGED = 1,
    KRYLOV_LIMIT = 2,
    KRYLOV_NONFINITE = 3,
    KRYLOV_BREAKDOWN = 4,
    KRYLOV_REFINE = 5, // mixed precision: a reliable update is due
};

// ----------------------------------------------------------------- SpMV layouts
// Node-block CSR: the tangent of a 3-dof-per-node mesh keeps its values in ordinary
// row-major CSR order (sorted columns), but every dof row of node n has the same column set
// {3m+j : m in adj(n), j < 3}, so the kernels read one int32 block column per 9 values:
//   row 3n+i, block s (node m = bcol[bptr[n]+s]), column 3m+j
//   -> val[9 bptr[n] + i (3 deg n) + 3 s + j].
struct nb_matrix
{
    int32_t nodes;
    const int32_t* bptr;
    const int32_t* bcol;
};

struct csr_matrix
{
    int32_t rows;
    const int32_t* rp;
    const int32_t* ci;
};

template <bool NT, typename T>
__device__ inline T ld(const T* p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// ROWS block rows per group of LPR lanes; one lane per 3x3 block (node pair): it reads the
// block's column index, the three x entries of the column node and the 9 values (3 runs of 3
// contiguous values, one per dof row), then the three row sums are reduced across the group.
// NT streams the values with nontemporal loads (read once per product) so that x stays cached.
// With DOT, also adds the lane's share of x . y: sum_i x_{n,i} (K_nm x_m)_i for its block, so
// that the shares of a row sum to x_n . y_n without waiting for the row reduction.
// b0, d: the rows' first block and block count (bptr[n], bptr[n+1] - bptr[n]; 0 for rows not ok).
template <int LPR, int ROWS, bool NT, typename VT, typename XT, typename AT>
__device__ inline void nb_rows(nb_matrix A, const VT* __restrict__ val, const XT* __restrict__ x,
                               const int32_t* n, const bool* ok, const int32_t* b0, const int32_t* d, int lane,
                               AT (*y)[3])
{
    AT s[ROWS][3];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) s[r][0] = s[r][1] = s[r][2] = AT(0);
    int32_t dmax = d[0];
#pragma unroll
    for (int r = 1; r < ROWS; ++r) dmax = d[r] > dmax ? d[r] : dmax;
    for (int q = lane; q < dmax; q += LPR)
    {
        VT v[ROWS][9];
        AT xs[ROWS][3];
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
        {
            bool const on = q < d[r];
            int64_t const m = on ? A.bcol[b0[r] + q] : 0;
            int64_t const rs = 3ll * d[r];
            const VT* p = val + 9ll * b0[r] + 3 * q;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) v[r][3 * i + j] = on ? ld<NT>(p + i * rs + j) : VT(0);
#pragma unroll
            for (int j = 0; j < 3; ++j) xs[r][j] = on ? AT(x[3 * m + j]) : AT(0);
        }
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
#pragma unroll
            for (int i = 0; i < 3; ++i)
                s[r][i] += AT(v[r][3 * i]) * xs[r][0] + AT(v[r][3 * i + 1]) * xs[r][1] + AT(v[r][3 * i + 2]) * xs[r][2];
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            AT t = s[r][i];
#pragma unroll
            for (int off = LPR / 2; off > 0; off >>= 1) t += __shfl_xor(t, off, LPR);
            y[r][i] = t;
        }
}

// y = A x (node-block layout); with DOT, also writes the block partial of x . y.
// MODE 0: y = Ax.  MODE 1: y = b - Ax (residual; b given in `bvec`).
// Node order: each XCD (workgroups b with b % 8 equal) owns one contiguous eighth of the
// nodes.  SWEEP=false: every workgroup walks its own contiguous chunk of that eighth.
// SWEEP=true: the workgroups of one XCD advance together through the eighth, workgroup j
// taking node groups j, j + W, j + 2W, ... (W workgroups per XCD), so the x entries that the
// concurrently running workgroups gather (neighbours within about one node plane) stay in
// that XCD's L2 instead of spanning the whole eighth.
template <int LPR, int ROWS, bool NT, bool SWEEP, typename VT, typename XT, typename YT, typename AT, bool DOT,
          int MODE>
__global__ __launch_bounds__(kBlock) void spmv_nb_kernel(nb_matrix A, const VT* __restrict__ val,
                                                         const XT* __restrict__ x, YT* __restrict__ y,
                                                         const double* __restrict__ bvec,
                                                         double* __restrict__ part,
                                                         const krylov_state* __restrict__ st, int gate)
{
    if (st && (st->done || (gate && !st->refine))) return;
    constexpr int G = kBlock / LPR; // lane groups per block
    int const g = threadIdx.x / LPR, l = threadIdx.x % LPR;
    int64_t start, end, stride;
    if (SWEEP && (gridDim.x & 7) == 0)
    {
        int64_t const per = (int64_t(A.nodes) + 7) / 8;
        int64_t const beg = int64_t(blockIdx.x & 7) * per;
        end = beg + per < A.nodes ? beg + per : A.nodes;
        stride = int64_t(gridDim.x >> 3) * G;
        start = beg + int64_t(blockIdx.x >> 3) * G + g;
    }
    else
    {
        int64_t beg;
        xcd_chunk(A.nodes, beg, end);
        stride = G;
        start = beg + g;
    }
    double dot = 0.0;
    for (int64_t n0 = start; n0 < end; n0 += stride * ROWS)
    {
        int32_t n[ROWS], b0[ROWS], d[ROWS];
        bool ok[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
        {
            ok[r] = n0 + r * stride < end;
            n[r] = ok[r] ? int32_t(n0 + r * stride) : 0;
            b0[r] = ok[r] ? A.bptr[n[r]] : 0;
            d[r] = ok[r] ? A.bptr[n[r] + 1] - b0[r] : 0;
        }
        // the row's own x (dot) and b (residual) entries, requested before the row's blocks so
        // that their latency overlaps (lanes >= 3 read a valid neighbour and drop it); the dot
        // uses the stored (rounded) y, consistent with the recurrence that uses it
        double br[ROWS];
        XT xr[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
        {
            int64_t const row = 3ll * n[r] + (l < 3 ? l : 2);
            br[r] = MODE == 1 ? bvec[row] : 0.0;
            xr[r] = DOT ? x[row] : XT(0);
        }
        AT yy[ROWS][3];
        nb_rows<LPR, ROWS, NT, VT, XT, AT>(A, val, x, n, ok, b0, d, l, yy);
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
        {
            // lane l < 3 takes row sum l (selects on named values: no indexed private array)
            AT const y0 = yy[r][0], y1 = yy[r][1], y2 = yy[r][2];
            AT yi = l == 1 ? y1 : y0;
            yi = l == 2 ? y2 : yi;
            YT const yo = MODE == 1 ? YT(br[r] - double(yi)) : YT(yi);
            bool const mine = ok[r] && l < 3;
            // every lane consumes its loads on every path (a load left pending on one path
            // would force a wait before the next row's loads reuse its register)
            if constexpr (DOT) dot += mine ? double(xr[r]) * double(yo) : 0.0;
            if (mine) y[3ll * n[r] + l] = yo;
        }
    }
    if constexpr (DOT)
    {
        dot = block_sum<kBlock>(dot);
        if (threadIdx.x == 0) part[blockIdx.x] = dot;
    }
}

// Sliced node-block layout: a copy of the node-block values in which every wavefront load
// reads 64 consecutive values.  Slice s holds the 64 consecutive nodes 64 s .. 64 s + 63
// (lane = node) and sptr[s+1] - sptr[s] steps, the largest block count among them.  Step kk of
// a node is its kk-th block (ascending column node, as in bcol); steps past the node's block
// count, and lanes past the last node, hold zero values with the node itself (node 0 past the
// end) as column.  Value (i, j) of step kk of slice s sits at
//   sval[((sptr[s] + kk) * 9 + 3 i + j) * 64 + lane],  its column node at scol[(sptr[s] + kk) * 64 + lane].
// scol16, when set, holds the same columns as 16-bit offsets from the lane's node (64 s + lane;
// lanes past the last node point at the last node), halving the column bytes.
struct sell_matrix
{
    int32_t nodes;
    int32_t slices;
    const int32_t* sptr;
    const int32_t* scol;
    const int16_t* scol16;
};

// y = A x on the sliced layout, one wavefront per slice, one node (three rows) per lane.  Each
// lane accumulates its three row sums over its steps in order, so no cross-lane reduction is
// needed.  With DOT, also writes the block partial of x . y (the stored, rounded y).
// MODE 0: y = Ax.  MODE 1: y = b - Ax.
// Slice order: XCD=false, wavefront w of workgroup b takes slices b W + w, stride gridDim W
// (consecutive slices on different XCDs, so every XCD's L2 fetches all of x).  XCD=true (a
// multiple of 8 workgroups): the workgroups b with b % 8 = e (one XCD) sweep the e-th eighth
// of the slices together, so each L2 fetches x only around its own eighth.
// NT: the values are read with nontemporal loads (streamed past the caches, which keep x).
template <int BS, bool XCD, bool I16, bool NT, typename VT, typename XT, typename YT, typename AT, bool DOT, int MODE>
__global__ __launch_bounds__(BS) void spmv_sell_kernel(sell_matrix A, const VT* __restrict__ val,
                                                       const XT* __restrict__ x, YT* __restrict__ y,
                                                       const double* __restrict__ bvec, double* __restrict__ part,
                                                       const krylov_state* __restrict__ st, int gate)
{
    if (st && (st->done || (gate && !st->refine))) return;
    constexpr int W = BS / 64;
    int const lane = threadIdx.x & 63;
    int32_t first, stride, last;
    if (XCD && (gridDim.x & 7) == 0)
    {
        int32_t const per = (A.slices + 7) / 8;
        int32_t const beg = int32_t(blockIdx.x & 7) * per;
        last = beg + per < A.slices ? beg + per : A.slices;
        first = beg + int32_t(blockIdx.x >> 3) * W + int32_t(threadIdx.x >> 6);
        stride = int32_t(gridDim.x >> 3) * W;
    }
    else
    {
        first = int32_t(blockIdx.x) * W + int32_t(threadIdx.x >> 6);
        stride = int32_t(gridDim.x) * W;
        last = A.slices;
    }
    double dot = 0.0;
    for (int32_t s = first; s < last; s += stride)
    {
        int64_t const n = int64_t(s) * 64 + lane;
        bool const ok = n < A.nodes;
        int64_t const row = 3 * (ok ? n : int64_t(A.nodes) - 1);
        // the node's own b (residual) and x (dot) entries, requested before its blocks
        double br0 = 0.0, br1 = 0.0, br2 = 0.0;
        XT xr0 = XT(0), xr1 = XT(0), xr2 = XT(0);
        if constexpr (MODE == 1)
        {
            br0 = bvec[row];
            br1 = bvec[row + 1];
            br2 = bvec[row + 2];
        }
        if constexpr (DOT)
        {
            xr0 = x[row];
            xr1 = x[row + 1];
            xr2 = x[row + 2];
        }
        int32_t const k0 = A.sptr[s], ks = A.sptr[s + 1] - k0;
        // FP32 accumulation: four accumulator sets take the steps in turn (steps kk, kk+4, ...
        // in one set) and are added pairwise at the end, so a row's rounding error grows like
        // that of a sum of ~7 terms instead of ~27 (the CSR-order kernels' lane tree gives
        // about the same); FP64 accumulation uses the first set only
        AT a0 = AT(0), a1 = AT(0), a2 = AT(0), b0 = AT(0), b1 = AT(0), b2 = AT(0);
        AT c0 = AT(0), c1 = AT(0), c2 = AT(0), d0 = AT(0), d1 = AT(0), d2 = AT(0);
        for (int32_t kk = 0; kk < ks; ++kk)
        {
            int64_t const blk = int64_t(k0) + kk;
            int64_t const m = I16 ? n + int64_t(A.scol16[blk * 64 + lane]) : int64_t(A.scol[blk * 64 + lane]);
            VT v[9];
#pragma unroll
            for (int t = 0; t < 9; ++t)
            {
                const VT* src = val + (blk * 9 + t) * 64 + lane;
                if constexpr (NT)
                    v[t] = __builtin_nontemporal_load(src);
                else
                    v[t] = *src;
            }
            AT const x0 = AT(x[3 * m]), x1 = AT(x[3 * m + 1]), x2 = AT(x[3 * m + 2]);
            AT const n0 = a0 + (AT(v[0]) * x0 + AT(v[1]) * x1 + AT(v[2]) * x2);
            AT const n1 = a1 + (AT(v[3]) * x0 + AT(v[4]) * x1 + AT(v[5]) * x2);
            AT const n2 = a2 + (AT(v[6]) * x0 + AT(v[7]) * x1 + AT(v[8]) * x2);
            if constexpr (sizeof(AT) == 4)
            {
                a0 = b0, a1 = b1, a2 = b2;
                b0 = c0, b1 = c1, b2 = c2;
                c0 = d0, c1 = d1, c2 = d2;
                d0 = n0, d1 = n1, d2 = n2;
            }
            else
            {
                // FP64 accumulation: one set (the others stay zero; fewer registers)
                a0 = n0, a1 = n1, a2 = n2;
            }
        }
        AT const r0 = (a0 + b0) + (c0 + d0), r1 = (a1 + b1) + (c1 + d1), r2 = (a2 + b2) + (c2 + d2);
        YT const y0 = MODE == 1 ? YT(br0 - double(r0)) : YT(r0);
        YT const y1 = MODE == 1 ? YT(br1 - double(r1)) : YT(r1);
        YT const y2 = MODE == 1 ? YT(br2 - double(r2)) : YT(r2);
        if (ok)
        {
            y[row] = y0;
            y[row + 1] = y1;
            y[row + 2] = y2;
        }
        if constexpr (DOT)
            dot += ok ? double(xr0) * double(y0) + double(xr1) * double(y1) + double(xr2) * double(y2) : 0.0;
    }
    if constexpr (DOT)
    {
        dot = block_sum<BS>(dot);
        if (threadIdx.x == 0) part[blockIdx.x] = dot;
    }
}

// Sliced copy of node-block values (layout above), with coalesced reads: one 256-thread
// workgroup per slice; for each dof row i the slice's 64 row segments (3 deg contiguous values
// each, CSR order val[9 bptr[n] + i (3 deg) + 3 kk + j]) are read into LDS (rows of `row_len`
// = 3 * max degree values), then written out 64 consecutive values at a time.  Steps past a
// node's block count get zeros.  Dynamic LDS: 64 * row_len values.
template <typename VT>
__global__ __launch_bounds__(256) void sell_pack_lds_kernel(sell_matrix A, const int32_t* __restrict__ bptr,
                                                            const VT* __restrict__ val, VT* __restrict__ sval,
                                                            int32_t row_len)
{
    extern __shared__ __align__(16) unsigned char pack_smem[];
    VT* sm = reinterpret_cast<VT*>(pack_smem);
    __shared__ int32_t sb0[64], sd[64];
    int32_t const s = blockIdx.x;
    int const tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid < 64)
    {
        int64_t const n = int64_t(s) * 64 + tid;
        bool const ok = n < A.nodes;
        sb0[tid] = ok ? bptr[n] : 0;
        sd[tid] = ok ? bptr[n + 1] - bptr[n] : 0;
    }
    int32_t const k0 = A.sptr[s], ks = A.sptr[s + 1] - k0;
    __syncthreads();
    for (int i = 0; i < 3; ++i)
    {
        for (int q = w; q < 64; q += 4)
        {
            int32_t const d = sd[q];
            const VT* src = val + 9ll * sb0[q] + int64_t(i) * 3 * d;
            for (int e = lane; e < 3 * d; e += 64) sm[q * row_len + e] = src[e];
        }
        __syncthreads();
        for (int c = w; c < 3 * ks; c += 4)
        {
            int const kk = c / 3, j = c - 3 * (c / 3);
            VT const v = kk < sd[lane] ? sm[lane * row_len + 3 * kk + j] : VT(0);
            sval[((int64_t(k0) + kk) * 9 + 3 * i + j) * 64 + lane] = v;
        }
        __syncthreads();
    }
}

// Node-block SpMV with the row pointers of 32 rows per load (32 lanes per block row, XCD sweep
// order as above, any row length).  Lane t of a lane group holds bptr of the group's row
// 32 c + t (and of the row after it) for the current chunk c, the next chunk's being requested
// a chunk ahead, and each row takes its first block and length by a shuffle instead of a
// dependent load: a row then waits on two memory round trips (columns with values, then the
// x gather) instead of three.
template <typename VT, typename XT, typename YT, typename AT, bool DOT, int MODE>
__global__ __launch_bounds__(kBlock, sizeof(VT) == 8 || sizeof(XT) == 8 ? 6 : 8) void spmv_nb_meta_kernel(nb_matrix A, const VT* __restrict__ val,
                                                              const XT* __restrict__ x, YT* __restrict__ y,
                                                              const double* __restrict__ bvec,
                                                              double* __restrict__ part,
                                                              const krylov_state* __restrict__ st, int gate)
{
    if (st && (st->done || (gate && !st->refine))) return;
    constexpr int LPR = 32;
    constexpr int G = kBlock / LPR;
    int const g = threadIdx.x / LPR, l = threadIdx.x % LPR;
    int64_t start, end, stride;
    if ((gridDim.x & 7) == 0)
    {
        int64_t const per = (int64_t(A.nodes) + 7) / 8;
        int64_t const beg = int64_t(blockIdx.x & 7) * per;
        end = beg + per < A.nodes ? beg + per : A.nodes;
        stride = int64_t(gridDim.x >> 3) * G;
        start = beg + int64_t(blockIdx.x >> 3) * G + g;
    }
    else
    {
        int64_t beg;
        xcd_chunk(A.nodes, beg, end);
        stride = G;
        start = beg + g;
    }
    int64_t const count = start < end ? (end - start + stride - 1) / stride : 0;
    // row pointers of row 32 c + l (clamped to the last row; past the end the row is empty)
    auto meta = [&](int64_t c, int32_t& lo, int32_t& hi) {
        int64_t const t = c * LPR + l;
        int64_t const tc = t < count ? t : (count > 0 ? count - 1 : 0);
        int64_t const n = count > 0 ? start + tc * stride : 0;
        lo = A.bptr[n];
        hi = t < count ? A.bptr[n + 1] : lo;
    };
    double dot = 0.0;
    int32_t mlo, mhi, nlo, nhi;
    meta(0, mlo, mhi);
    meta(1, nlo, nhi);
    for (int64_t t = 0; t < count; ++t)
    {
        int const ln = int(t % LPR);
        if (ln == 0 && t > 0)
        {
            mlo = nlo;
            mhi = nhi;
            meta(t / LPR + 1, nlo, nhi);
        }
        int32_t const b0 = __shfl(mlo, ln, LPR);
        int32_t const d = __shfl(mhi, ln, LPR) - b0;
        int32_t const n = int32_t(start + t * stride);
        int64_t const row = 3ll * n + (l < 3 ? l : 2);
        double const br = MODE == 1 ? bvec[row] : 0.0;
        XT const xr = DOT ? x[row] : XT(0);
        bool const ok = true;
        AT yy[1][3];
        nb_rows<LPR, 1, false, VT, XT, AT>(A, val, x, &n, &ok, &b0, &d, l, yy);
        AT yi = l == 1 ? yy[0][1] : yy[0][0];
        yi = l == 2 ? yy[0][2] : yi;
        YT const yo = MODE == 1 ? YT(br - double(yi)) : YT(yi);
        if constexpr (DOT) dot += l < 3 ? double(xr) * double(yo) : 0.0;
        if (l < 3) y[3ll * n + l] = yo;
    }
    if constexpr (DOT)
    {
        dot = block_sum<kBlock>(dot);
        if (threadIdx.x == 0) part[blockIdx.x] = dot;
    }
}

// Software-pipelined node-block SpMV (32 lanes per block row, at most 32 blocks per row,
// XCD sweep order as above).  Iteration t of a lane group gathers x for row t, requests the
// row pointers of row t+2 and the column indices and 9 values per lane of row t+1, and only
// then combines row t, so each iteration exposes one memory latency with the next row's
// loads in flight.  Every load is unconditional from a clamped, valid address and masked
// afterwards: a load under a branch would force a wait for all outstanding loads at the join.
template <bool NT, typename VT, typename XT, typename YT, typename AT, bool DOT, int MODE>
__global__ __launch_bounds__(kBlock) void spmv_nb_pipe_kernel(nb_matrix A, const VT* __restrict__ val,
                                                              const XT* __restrict__ x, YT* __restrict__ y,
                                                              const double* __restrict__ bvec,
                                                              double* __restrict__ part,
                                                              const krylov_state* __restrict__ st, int gate)
{
    if (st && (st->done || (gate && !st->refine))) return;
    constexpr int LPR = 32;
    constexpr int G = kBlock / LPR;
    int const g = threadIdx.x / LPR, l = threadIdx.x % LPR;
    int64_t start, end, stride;
    if ((gridDim.x & 7) == 0)
    {
        int64_t const per = (int64_t(A.nodes) + 7) / 8;
        int64_t const beg = int64_t(blockIdx.x & 7) * per;
        end = beg + per < A.nodes ? beg + per : A.nodes;
        stride = int64_t(gridDim.x >> 3) * G;
        start = beg + int64_t(blockIdx.x >> 3) * G + g;
    }
    else
    {
        int64_t beg;
        xcd_chunk(A.nodes, beg, end);
        stride = G;
        start = beg + g;
    }
    int64_t const count = start < end ? (end - start + stride - 1) / stride : 0;
    double dot = 0.0;
    if (count > 0)
    {
        // row pointers of row t (clamped to the last row; degree 0 past the end)
        auto meta = [&](int64_t t, int32_t& b0, int32_t& d) {
            int64_t const n = start + (t < count ? t : count - 1) * stride;
            int32_t const lo = A.bptr[n], hi = A.bptr[n + 1];
            b0 = lo;
            d = t < count ? hi - lo : 0;
        };
        // lane l's block of a row: column index and 9 values; lanes past the row re-read its
        // first block (cached) and are masked through their x entries at use
        auto fetch = [&](int32_t b0, int32_t d, VT* v, int32_t& col, bool& on) {
            on = l < d;
            int const q = on ? l : 0;
            int64_t const rs = 3ll * (d > 0 ? d : 1);
            col = A.bcol[b0 + q];
            __builtin_amdgcn_sched_barrier(0);
            const VT* p = val + 9ll * b0 + 3 * q;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) v[3 * i + j] = ld<NT>(p + i * rs + j);
        };
        int32_t b0n, dn; // row pointers of the next row to fetch
        // one pipeline stage: combine row t (values vc, index cc) while row t+1 is fetched into vn
        auto stage = [&](int64_t t, const VT* vc, int32_t cc, bool oc, VT* vn, int32_t& cn, bool& on) {
            AT xs[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) xs[j] = AT(x[3ll * cc + j]);
            int32_t b0nn, dnn;
            meta(t + 2, b0nn, dnn);
            // issue order x, pointers, indices, values: the next stage then waits only for the
            // indices while the values are still in flight
            __builtin_amdgcn_sched_barrier(0);
            fetch(b0n, dn, vn, cn, on);
            int64_t const row0 = 3 * (start + t * stride);
            double xr = 0.0, br = 0.0;
            if constexpr (DOT) xr = double(x[row0 + (l < 3 ? l : 2)]);
            if constexpr (MODE == 1) br = bvec[row0 + (l < 3 ? l : 2)];
#pragma unroll
            for (int j = 0; j < 3; ++j) xs[j] = oc ? xs[j] : AT(0);
            AT s0 = AT(vc[0]) * xs[0] + AT(vc[1]) * xs[1] + AT(vc[2]) * xs[2];
            AT s1 = AT(vc[3]) * xs[0] + AT(vc[4]) * xs[1] + AT(vc[5]) * xs[2];
            AT s2 = AT(vc[6]) * xs[0] + AT(vc[7]) * xs[1] + AT(vc[8]) * xs[2];
#pragma unroll
            for (int off = LPR / 2; off > 0; off >>= 1)
            {
                s0 += __shfl_xor(s0, off, LPR);
                s1 += __shfl_xor(s1, off, LPR);
                s2 += __shfl_xor(s2, off, LPR);
            }
            AT const yi = l == 0 ? s0 : (l == 1 ? s1 : s2);
            YT const yo = MODE == 1 ? YT(br - double(yi)) : YT(yi);
            if (l < 3) y[row0 + l] = yo;
            if constexpr (DOT) dot += l < 3 ? xr * double(yo) : 0.0;
            b0n = b0nn;
            dn = dnn;
        };
        int32_t b0, d;
        meta(0, b0, d);
        meta(1, b0n, dn);
        // two register sets, alternated by a 2x unrolled loop, so no loaded value is copied
        VT va[9], vb[9];
        int32_t ca, cb;
        bool oa, ob;
        fetch(b0, d, va, ca, oa);
        int64_t t = 0;
        for (; t + 1 < count; t += 2)
        {
            stage(t, va, ca, oa, vb, cb, ob);
            stage(t + 1, vb, cb, ob, va, ca, oa);
        }
        if (t < count) stage(t, va, ca, oa, vb, cb, ob);
    }
    if constexpr (DOT)
    {
        dot = block_sum<kBlock>(dot);
        if (threadIdx.x == 0) part[blockIdx.x] = dot;
    }
}

// LDS-staged node-block SpMV (32 lanes per block row, at most 32 blocks per row, XCD sweep
// order).  The 8 node rows that one workgroup combines per iteration are consecutive, so their
// values form one contiguous segment of the value array; the workgroup streams it into LDS
// with 16-byte loads (every load instruction fully used), then each lane takes its block's
// 9 values from LDS, gathers x and the row sums reduce across the 32 lanes as in
// spmv_nb_kernel.  The value array must extend 16 bytes past nnz (loads round the segment end
// up to a whole vector).
template <typename VT, typename XT, typename YT, typename AT, bool DOT, int MODE>
__global__ __launch_bounds__(kBlock) void spmv_nb_lds_kernel(nb_matrix A, const VT* __restrict__ val,
                                                             const XT* __restrict__ x, YT* __restrict__ y,
                                                             const double* __restrict__ bvec,
                                                             double* __restrict__ part,
                                                             const krylov_state* __restrict__ st, int gate)
{
    if (st && (st->done || (gate && !st->refine))) return;
    constexpr int LPR = 32;
    constexpr int G = kBlock / LPR;              // node rows per iteration
    constexpr int VPV = 16 / int(sizeof(VT));    // values per 16-byte vector
    constexpr int CAP = G * 9 * LPR + 2 * VPV;   // segment of G rows of <= 32 blocks + alignment
    typedef VT vec __attribute__((ext_vector_type(VPV)));
    __shared__ __align__(16) VT sv[CAP];
    int const g = threadIdx.x / LPR, l = threadIdx.x % LPR;
    int64_t const per = (int64_t(A.nodes) + 7) / 8;
    int64_t beg, end, step, first;
    if ((gridDim.x & 7) == 0)
    {
        beg = int64_t(blockIdx.x & 7) * per;
        end = beg + per < A.nodes ? beg + per : A.nodes;
        step = int64_t(gridDim.x >> 3) * G;
        first = beg + int64_t(blockIdx.x >> 3) * G;
    }
    else
    {
        xcd_chunk(A.nodes, beg, end);
        step = G;
        first = beg;
    }
    double dot = 0.0;
    for (int64_t n0 = first; n0 < end; n0 += step)
    {
        int64_t const rows = end - n0 < G ? end - n0 : G;
        int64_t const s0 = 9ll * A.bptr[n0], s1 = 9ll * A.bptr[n0 + rows];
        int64_t const a0 = s0 & ~int64_t(VPV - 1);
        int const nv = int((s1 - a0 + VPV - 1) / VPV);
        const vec* src = reinterpret_cast<const vec*>(val + a0);
        vec* dst = reinterpret_cast<vec*>(sv);
        for (int v = threadIdx.x; v < nv; v += kBlock) dst[v] = __builtin_nontemporal_load(src + v);
        __syncthreads();
        if (g < rows)
        {
            int64_t const n = n0 + g;
            int32_t const b0 = A.bptr[n], d = A.bptr[n + 1] - b0;
            bool const on = l < d;
            int const q = on ? l : 0;
            int64_t const m = A.bcol[b0 + q];
            AT xs[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) xs[j] = on ? AT(x[3 * m + j]) : AT(0);
            const VT* pv = sv + (9ll * b0 - a0) + 3 * q;
            int const rs = 3 * d;
            AT s0v = AT(pv[0]) * xs[0] + AT(pv[1]) * xs[1] + AT(pv[2]) * xs[2];
            AT s1v = AT(pv[rs]) * xs[0] + AT(pv[rs + 1]) * xs[1] + AT(pv[rs + 2]) * xs[2];
            AT s2v = AT(pv[2 * rs]) * xs[0] + AT(pv[2 * rs + 1]) * xs[1] + AT(pv[2 * rs + 2]) * xs[2];
#pragma unroll
            for (int off = LPR / 2; off > 0; off >>= 1)
            {
                s0v += __shfl_xor(s0v, off, LPR);
                s1v += __shfl_xor(s1v, off, LPR);
                s2v += __shfl_xor(s2v, off, LPR);
            }
            if (l < 3)
            {
                AT const yi = l == 0 ? s0v : (l == 1 ? s1v : s2v);
                int64_t const row = 3 * n + l;
                YT yo;
                if constexpr (MODE == 1)
                    yo = YT(bvec[row] - double(yi));
                else
                    yo = YT(yi);
                y[row] = yo;
                if constexpr (DOT) dot += double(x[row]) * double(yo);
            }
        }
        __syncthreads();
    }
    if constexpr (DOT)
    {
        dot = block_sum<kBlock>(dot);
        if (threadIdx.x == 0) part[blockIdx.x] = dot;
    }
}

// Streaming node-block SpMV (32 lanes per block row, at most 32 blocks per row, XCD sweep
// order).  Each iteration of a workgroup combines 8 consecutive node rows, whose values, block
// columns and row pointers are contiguous segments of their arrays.  The workgroup keeps the
// current rows' segments in LDS and, before combining them, requests the next rows' segments
// into registers with 16-byte loads (and the row pointers of the rows after), so the value
// stream never waits on the x gather or the reductions of the rows being combined.  The value
// array must extend one 16-byte vector past nnz.
template <typename VT, typename XT, typename YT, typename AT, bool DOT, int MODE>
__global__ __launch_bounds__(kBlock) void spmv_nb_stream_kernel(nb_matrix A, const VT* __restrict__ val,
                                                                const XT* __restrict__ x, YT* __restrict__ y,
                                                                const double* __restrict__ bvec,
                                                                double* __restrict__ part,
                                                                const krylov_state* __restrict__ st, int gate)
{
    if (st && (st->done || (gate && !st->refine))) return;
    constexpr int LPR = 32;
    constexpr int G = kBlock / LPR;                   // node rows per iteration
    constexpr int VPV = 16 / int(sizeof(VT));         // values per 16-byte vector
    constexpr int NV = (G * 9 * LPR + 2 * VPV) / VPV; // vectors of a segment incl. alignment
    constexpr int PER = (NV + kBlock - 1) / kBlock;   // vectors per thread
    typedef VT vec __attribute__((ext_vector_type(VPV)));
    __shared__ __align__(16) VT sv[NV * VPV];
    __shared__ int32_t sc[G * LPR];
    __shared__ int32_t sp[2][G + 1];
    int const t = threadIdx.x;
    int const g = t / LPR, l = t % LPR;
    int64_t beg, end, step, first;
    if ((gridDim.x & 7) == 0)
    {
        int64_t const per = (int64_t(A.nodes) + 7) / 8;
        beg = int64_t(blockIdx.x & 7) * per;
        end = beg + per < A.nodes ? beg + per : A.nodes;
        step = int64_t(gridDim.x >> 3) * G;
        first = beg + int64_t(blockIdx.x >> 3) * G;
    }
    else
    {
        xcd_chunk(A.nodes, beg, end);
        step = G;
        first = beg;
    }
    int64_t const iters = first < end ? (end - first + step - 1) / step : 0;
    double dot = 0.0;
    // row pointers bptr[n0 + k], k <= G, of iteration it (clamped at the end of the range)
    auto pointer = [&](int64_t it) -> int32_t {
        int64_t const n0 = first + it * step;
        int64_t const rows = end - n0 < G ? end - n0 : G;
        return A.bptr[n0 + (t < rows ? t : rows)];
    };
    vec pv[PER];
    int32_t pc = 0, pp = 0;
    // requests of the segments of iteration it, whose row pointers are in sp[it & 1]
    auto request = [&](int64_t it) {
        int32_t const* q = sp[it & 1];
        int64_t const s0 = 9ll * q[0], s1 = 9ll * q[G];
        int64_t const a0 = s0 & ~int64_t(VPV - 1);
        int const nv = int((s1 - a0 + VPV - 1) / VPV);
        const vec* src = reinterpret_cast<const vec*>(val + a0);
#pragma unroll
        for (int k = 0; k < PER; ++k)
        {
            int const v = t + k * kBlock;
            int const vc = v < nv ? v : 0;
            pv[k] = __builtin_nontemporal_load(src + vc);
        }
        int const nc = q[G] - q[0];
        pc = A.bcol[q[0] + (t < nc ? t : 0)];
    };
    auto store = [&](int64_t it) {
        int32_t const* q = sp[it & 1];
        int64_t const s0 = 9ll * q[0], s1 = 9ll * q[G];
        int64_t const a0 = s0 & ~int64_t(VPV - 1);
        int const nv = int((s1 - a0 + VPV - 1) / VPV);
        vec* dst = reinterpret_cast<vec*>(sv);
#pragma unroll
        for (int k = 0; k < PER; ++k)
        {
            int const v = t + k * kBlock;
            if (v < nv) dst[v] = pv[k];
        }
        if (t < q[G] - q[0]) sc[t] = pc;
    };
    if (iters > 0)
    {
        if (t <= G) sp[0][t] = pointer(0);
        if (t <= G && iters > 1) sp[1][t] = pointer(1);
        __syncthreads();
        request(0);
        store(0);
        __syncthreads();
        for (int64_t it = 0; it < iters; ++it)
        {
            bool const more = it + 1 < iters;
            if (more) request(it + 1);
            if (t <= G && it + 2 < iters) pp = pointer(it + 2);
            // combine the current rows from LDS
            int64_t const n0 = first + it * step;
            int64_t const rows = end - n0 < G ? end - n0 : G;
            int32_t const* q = sp[it & 1];
            if (g < rows)
            {
                int64_t const n = n0 + g;
                int32_t const b0 = q[g], d = q[g + 1] - b0;
                bool const on = l < d;
                int const qq = on ? l : 0;
                int64_t const m = sc[b0 - q[0] + qq];
                int64_t const row = 3 * n + (l < 3 ? l : 2);
                double const xr = DOT ? double(x[row]) : 0.0;
                double const br = MODE == 1 ? bvec[row] : 0.0;
                AT xs[3];
#pragma unroll
                for (int j = 0; j < 3; ++j) xs[j] = on ? AT(x[3 * m + j]) : AT(0);
                int64_t const a0 = (9ll * q[0]) & ~int64_t(VPV - 1);
                const VT* pvv = sv + (9ll * b0 - a0) + 3 * qq;
                int const rs = 3 * d;
                AT s0v = AT(pvv[0]) * xs[0] + AT(pvv[1]) * xs[1] + AT(pvv[2]) * xs[2];
                AT s1v = AT(pvv[rs]) * xs[0] + AT(pvv[rs + 1]) * xs[1] + AT(pvv[rs + 2]) * xs[2];
                AT s2v = AT(pvv[2 * rs]) * xs[0] + AT(pvv[2 * rs + 1]) * xs[1] + AT(pvv[2 * rs + 2]) * xs[2];
#pragma unroll
                for (int off = LPR / 2; off > 0; off >>= 1)
                {
                    s0v += __shfl_xor(s0v, off, LPR);
                    s1v += __shfl_xor(s1v, off, LPR);
                    s2v += __shfl_xor(s2v, off, LPR);
                }
                if (l < 3)
                {
                    AT const yi = l == 0 ? s0v : (l == 1 ? s1v : s2v);
                    YT const yo = MODE == 1 ? YT(br - double(yi)) : YT(yi);
                    y[3 * n + l] = yo;
                    if constexpr (DOT) dot += xr * double(yo);
                }
            }
            __syncthreads();
            if (more) store(it + 1);
            if (t <= G && it + 2 < iters) sp[it & 1][t] = pp;
            __syncthreads();
        }
    }
    if constexpr (DOT)
    {
        dot = block_sum<kBlock>(dot);
        if (threadIdx.x == 0) part[blockIdx.x] = dot;
    }
}

// streaming read of n doubles in 16-byte loads (calibration of the achievable read bandwidth)
__global__ __launch_bounds__(kBlock) void stream_sum_kernel(int64_t n, const double* __restrict__ v,
                                                            double* __restrict__ out)
{
    int64_t beg, end;
    xcd_chunk(n / 2, beg, end);
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2* v2 = reinterpret_cast<const d2*>(v);
    double s = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock)
    {
        d2 const t = __builtin_nontemporal_load(v2 + i);
        s += t.x + t.y;
    }
    s = block_sum<kBlock>(s);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// copy of n values (calibration of the FETCH_SIZE / WRITE_SIZE counters)
template <typename VT>
__global__ __launch_bounds__(kBlock) void stream_copy_kernel(int64_t n, const VT* __restrict__ a, VT* __restrict__ b)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// Generic CSR (any sparsity): LPR lanes per row.
template <int LPR, typename VT, typename XT, typename YT, typename AT, bool DOT, int MODE>
__global__ __launch_bounds__(kBlock) void spmv_csr_kernel(csr_matrix A, const VT* __restrict__ val,
                                                          const XT* __restrict__ x, YT* __restrict__ y,
                                                          const double* __restrict__ bvec,
                                                          double* __restrict__ part,
                                                          const krylov_state* __restrict__ st, int gate)
{
    if (st && (st->done || (gate && !st->refine))) return;
    int64_t beg, end;
    xcd_chunk(A.rows, beg, end);
    int const g = threadIdx.x / LPR, l = threadIdx.x % LPR;
    double dot = 0.0;
    for (int64_t r = beg + g; r < end; r += kBlock / LPR)
    {
        AT s = 0;
        for (int32_t p = A.rp[r] + l; p < A.rp[r + 1]; p += LPR) s += AT(val[p]) * AT(x[A.ci[p]]);
#pragma unroll
        for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LPR);
        if (l == 0)
        {
            if constexpr (MODE == 1)
                y[r] = YT(bvec[r] - double(s));
            else
                y[r] = YT(s);
            if constexpr (DOT) dot += double(x[r]) * double(YT(s));
        }
    }
    if constexpr (DOT)
    {
        dot = block_sum<kBlock>(dot);
        if (threadIdx.x == 0) part[blockIdx.x] = dot;
    }
}

// Jacobi inverse from the CSR diagonal.  Eigen's DiagonalPreconditioner uses 1 where the
// diagonal is zero or missing; the CUDA solver inverts blindly (cwiseInverse).
template <typename T>
__global__ __launch_bounds__(kBlock) void jacobi_kernel(csr_matrix A, const double* __restrict__ val,
                                                        T* __restrict__ minv, int eigen_rule)
{
    int64_t const r = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (r >= A.rows) return;
    int32_t lo = A.rp[r], hi = A.rp[r + 1] - 1;
    double d = 0.0;
    bool found = false;
    while (lo <= hi)
    {
        int32_t const mid = (lo + hi) >> 1;
        int32_t const c = A.ci[mid];
        if (c == r)
        {
            d = val[mid];
            found = true;
            break;
        }
        if (c < r)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    double inv;
    if (eigen_rule)
        inv = (found && d != 0.0) ? 1.0 / d : 1.0;
    else
        inv = 1.0 / d;
    minv[r] = T(inv);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void jacobi32_kernel(csr_matrix A, const float* __restrict__ val,
                                                          T* __restrict__ minv, int eigen_rule)
{
    int64_t const r = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (r >= A.rows) return;
    double d = 0.0;
    bool found = false;
    for (int32_t p = A.rp[r]; p < A.rp[r + 1]; ++p)
        if (A.ci[p] == r)
        {
            d = double(val[p]);
            found = true;
            break;
        }
    double const inv = eigen_rule ? ((found && d != 0.0) ? 1.0 / d : 1.0) : 1.0 / d;
    minv[r] = T(inv);
}

// ------------------------------------------------------------------ CG (FP64)
// init: x = 0, r = b, p = M^-1 r; partials of r.r and r.z
__global__ __launch_bounds__(kBlock) void cg_init_kernel(int64_t n, double* __restrict__ x, double* __restrict__ r,
                                                         double* __restrict__ p, const double* __restrict__ b,
                                                         const double* __restrict__ minv,
                                                         double* __restrict__ part_rr,
                                                         double* __restrict__ part_rz)
{
    int64_t beg, end;
    xcd_chunk(n, beg, end);
    double rr = 0.0, rz = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock)
    {
        double const ri = b[i];
        double const zi = minv[i] * ri;
        x[i] = 0.0;
        r[i] = ri;
        p[i] = zi;
        rr += ri * ri;
        rz += ri * zi;
    }
    rr = block_sum<kBlock>(rr);
    rz = block_sum<kBlock>(rz);
    if (threadIdx.x == 0)
    {
        part_rr[blockIdx.x] = rr;
        part_rz[blockIdx.x] = rz;
    }
}

__global__ __launch_bounds__(kBlock) void cg_init_state_kernel(const double* __restrict__ part_rr,
                                                               const double* __restrict__ part_rz, int npart,
                                                               krylov_state* __restrict__ st, int criterion,
                                                               double tol, int max_it)
{
    double const rr = reduce_partials<kBlock>(part_rr, npart);
    double const rz = reduce_partials<kBlock>(part_rz, npart);
    if (threadIdx.x != 0) return;
    krylov_state s{};
    s.rz = rz;
    s.rr = rr;
    s.rhs2 = rr;
    s.it = 0;
    s.done = KRYLOV_RUNNING;
    if (criterion == NEON_STOP_RELATIVE)
    {
        s.threshold = fmax(tol * tol * rr, DBL_MIN);
        if (rr == 0.0 || rr < s.threshold) s.done = KRYLOV_CONVERGED;
    }
    else
    {
        s.threshold = 0.0;
        if (!(sqrt(rz) > tol)) s.done = KRYLOV_CONVERGED;
    }
    if (s.done == KRYLOV_RUNNING && max_it <= 0) s.done = KRYLOV_LIMIT;
    if (!isfinite(rr) || !isfinite(rz)) s.done = KRYLOV_NONFINITE;
    *st = s;
}

// x += alpha p, r -= alpha Ap with alpha = rz / (p.Ap); partials of r.r and r.z
// The x/r updates read and write x and read Ap with nontemporal accesses: neither is needed by
// the next product, so the caches keep p (the vector the product gathers) instead (FP64 CG
// iteration 917 -> 887 us at 6.4 M dofs; the FP32 recurrence is unchanged).
constexpr bool kVectorNT = true;
template <bool NT, typename T>
__device__ inline T load_stream(const T* p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}
template <bool NT, typename T>
__device__ inline void store_stream(T* p, T v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

__global__ __launch_bounds__(kBlock) void cg_update_xr_kernel(int64_t n, double* __restrict__ x,
                                                              double* __restrict__ r,
                                                              const double* __restrict__ p,
                                                              const double* __restrict__ Ap,
                                                              const double* __restrict__ minv,
                                                              const double* __restrict__ part_pAp, int npart,
                                                              const krylov_state* __restrict__ st,
                                                              double* __restrict__ part_rr,
                                                              double* __restrict__ part_rz)
{
    if (st->done) return;
    double const pAp = reduce_partials<kBlock>(part_pAp, npart);
    double const alpha = st->rz / pAp;
    int64_t beg, end;
    xcd_chunk_aligned<2>(n, beg, end);
    double rr = 0.0, rz = 0.0;
    auto body = [&](double& xi, double& ri, double pi, double api, double mi) {
        xi += alpha * pi;
        ri = ri - alpha * api;
        double const zi = mi * ri;
        rr += ri * ri;
        rz += ri * zi;
    };
    // two entries per thread in 16-byte accesses, then the chunk's odd tail
    int64_t const full = beg + (end - beg) / 2 * 2;
    for (int64_t i = beg + 2 * threadIdx.x; i < full; i += 2 * kBlock)
    {
        double2v xv = load_stream<kVectorNT>(reinterpret_cast<const double2v*>(x + i)),
                 rv = *reinterpret_cast<const double2v*>(r + i);
        double2v const pv = *reinterpret_cast<const double2v*>(p + i),
                       av = load_stream<kVectorNT>(reinterpret_cast<const double2v*>(Ap + i)),
                       mv = *reinterpret_cast<const double2v*>(minv + i);
        double x0 = xv.x, x1 = xv.y, r0 = rv.x, r1 = rv.y;
        body(x0, r0, pv.x, av.x, mv.x);
        body(x1, r1, pv.y, av.y, mv.y);
        store_stream<kVectorNT>(reinterpret_cast<double2v*>(x + i), double2v{x0, x1});
        *reinterpret_cast<double2v*>(r + i) = double2v{r0, r1};
    }
    for (int64_t i = full + threadIdx.x; i < end; i += kBlock)
    {
        double xi = x[i], ri = r[i];
        body(xi, ri, p[i], Ap[i], minv[i]);
        x[i] = xi;
        r[i] = ri;
    }
    rr = block_sum<kBlock>(rr);
    rz = block_sum<kBlock>(rz);
    if (threadIdx.x == 0)
    {
        part_rr[blockIdx.x] = rr;
        part_rz[blockIdx.x] = rz;
    }
}

// Convergence decision, beta, p = z + beta p; block 0 writes the next scalar state.
__global__ __launch_bounds__(kBlock) void cg_update_p_kernel(int64_t n, double* __restrict__ p,
                                                             const double* __restrict__ r,
                                                             const double* __restrict__ minv,
                                                             const double* __restrict__ part_rr,
                                                             const double* __restrict__ part_rz, int npart,
                                                             const krylov_state* __restrict__ st,
                                                             krylov_state* __restrict__ st_next, int criterion,
                                                             double tol, int max_it)
{
    krylov_state s = *st;
    if (s.done)
    {
        if (blockIdx.x == 0 && threadIdx.x == 0) *st_next = s;
        return;
    }
    double const rr = reduce_partials<kBlock>(part_rr, npart);
    double const rz = reduce_partials<kBlock>(part_rz, npart);
    int32_t it_new;
    int32_t done = KRYLOV_RUNNING;
    if (criterion == NEON_STOP_RELATIVE)
    {
        bool const conv = rr < s.threshold; // Eigen breaks before counting the iteration
        it_new = conv ? s.it : s.it + 1;
        if (conv)
            done = KRYLOV_CONVERGED;
        else if (it_new >= max_it)
            done = KRYLOV_LIMIT;
    }
    else
    {
        it_new = s.it + 1; // the CUDA loop counts at the top
        if (it_new >= max_it)
            done = KRYLOV_LIMIT;
        else if (!(sqrt(rz) > tol))
            done = KRYLOV_CONVERGED;
    }
    if (!isfinite(rr) || !isfinite(rz)) done = KRYLOV_NONFINITE;
    double const beta = rz / s.rz;
    if (done == KRYLOV_RUNNING)
    {
        int64_t beg, end;
        xcd_chunk_aligned<2>(n, beg, end);
        int64_t const full = beg + (end - beg) / 2 * 2;
        for (int64_t i = beg + 2 * threadIdx.x; i < full; i += 2 * kBlock)
        {
            double2v const mv = *reinterpret_cast<const double2v*>(minv + i),
                           rv = *reinterpret_cast<const double2v*>(r + i),
                           pv = *reinterpret_cast<const double2v*>(p + i);
            *reinterpret_cast<double2v*>(p + i) = double2v{mv.x * rv.x + beta * pv.x, mv.y * rv.y + beta * pv.y};
        }
        for (int64_t i = full + threadIdx.x; i < end; i += kBlock) p[i] = minv[i] * r[i] + beta * p[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        s.rr = rr;
        s.rz = rz;
        s.beta = beta;
        s.it = it_new;
        s.done = done;
        *st_next = s;
    }
}

// --------------------------------------------- fused CG vector update (one launch)
// Both vector phases of a CG iteration in one cooperative launch, with a grid barrier between
// them for r.z:
//   phase 1: alpha = rz / p.Ap;  x += alpha p;  r -= alpha Ap;  z = M^-1 r;  partials of r.r, r.z
//   phase 2: convergence decision and beta = rz_new / rz;  p = z + beta p.
// Thread g of the grid owns the 16-byte vectors g, g + T, g + 2T, ... (T threads in all) in
// both phases and keeps z (and, with KEEP_P, the old p) of its vectors in registers across
// the barrier, so phase 2 writes p without reading minv and r again (and, with KEEP_P, without
// reading p): per iteration 8 (KEEP_P) or 9 vector passes instead of the two kernels' 11.
// The grid is at most the number of co-resident workgroups (cooperative launch), and
// EPT entries per thread cover n.  MIXED: the FP32 recurrence of the mixed-precision CG on an
// iteration without a reliable-update test (no convergence decision here, as in
// mixed_update_p_kernel with decided = 0); otherwise the FP64 CG with `criterion`.
template <typename T>
struct vec16_of;
template <>
struct vec16_of<double>
{
    typedef double2v type;
    static constexpr int width = 2;
};
template <>
struct vec16_of<float>
{
    typedef float4v type;
    static constexpr int width = 4;
};

// Waves per SIMD the fused update is compiled for (register budget 512 / waves): with EPT
// 8 / 16 / 32 entries per thread, 8 / 7 / 4 waves, so that the grid of n / (256 EPT)
// workgroups fits co-resident on 256 compute units up to 8.4 M entries (EPT 32).
template <int EPT>
constexpr int fused_waves()
{
    return EPT <= 8 ? 8 : EPT <= 16 ? 7 : 4;
}

// Grid barrier of a cooperative launch (every workgroup co-resident).  bar[0] counts arrivals:
// workgroup 0 adds 2^31 - (nblocks - 1), the others 1, so the word's top bit flips exactly
// when the last workgroup arrives and its low bits return to their value (the word needs no
// reset between barriers).  Each workgroup's writes before the barrier are released by an
// agent-scope fence and acquired after it.  The wait is bounded (a poll is a coherent load of
// about a microsecond; a barrier normally completes within tens of microseconds): after
// kBarrierSpins polls the workgroup sets bar[1] and goes on; the fused update then ends the
// solve as non-finite and the host reports the time-out.
constexpr uint32_t kBarrierSpins = 1u << 17;
__device__ inline void grid_barrier(uint32_t* bar)
{
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t const add = blockIdx.x == 0 ? 0x80000000u - (gridDim.x - 1u) : 1u;
        __threadfence();
        uint32_t const old = __hip_atomic_fetch_add(bar, add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t spins = 0;
        while (((old ^ __hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) & 0x80000000u) == 0u)
        {
            __builtin_amdgcn_s_sleep(1);
            if (++spins == kBarrierSpins)
            {
                __hip_atomic_store(bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
        __threadfence();
    }
    __syncthreads();
}

// f(std::integral_constant<int, 0>{}), ..., f(std::integral_constant<int, K - 1>{}): a loop
// whose index is a constant expression in every copy, so arrays indexed by it stay in registers
template <typename F, int... I>
__device__ inline void static_for_impl(F&& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int K, typename F>
__device__ inline void static_for(F&& f)
{
    static_for_impl(f, std::make_integer_sequence<int, K>{});
}

template <typename T, int EPT, bool KEEP_P, bool MIXED>
__global__ __launch_bounds__(kBlock, fused_waves<EPT>()) void cg_fused_update_kernel(
    int64_t n, T* __restrict__ x, T* __restrict__ r, T* __restrict__ p, const T* __restrict__ Ap,
    const T* __restrict__ minv, const double* __restrict__ part_pAp, int npart, const krylov_state* __restrict__ st,
    krylov_state* __restrict__ st_next, double* __restrict__ part_rr, double* __restrict__ part_rz, int criterion,
    double tol, int max_it, uint32_t* __restrict__ bar)
{
    typedef typename vec16_of<T>::type V;
    constexpr int VW = vec16_of<T>::width;
    constexpr int K = EPT / VW;
    static_assert(EPT % VW == 0, "entries per thread: whole 16-byte vectors");
    krylov_state s = *st;
    if (s.done)
    {
        if (blockIdx.x == 0 && threadIdx.x == 0) *st_next = s;
        return;
    }
    double const pAp = reduce_partials<kBlock>(part_pAp, npart);
    T const alpha = T(s.rz / pAp);
    int64_t const total = int64_t(gridDim.x) * kBlock;
    int64_t const g = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    int64_t const full = n / VW; // whole vectors; the last thread of the grid takes the rest
    bool const tail = g == total - 1 && full * VW < n;
    V zk[K], pk[KEEP_P ? K : 1];
    double rr = 0.0, rz = 0.0;
    static_for<K>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        int64_t const v = g + k * total;
        if (v < full)
        {
            V xv = load_stream<kVectorNT>(reinterpret_cast<const V*>(x) + v);
            V rv = reinterpret_cast<const V*>(r)[v];
            V const pv = reinterpret_cast<const V*>(p)[v];
            V const av = load_stream<kVectorNT>(reinterpret_cast<const V*>(Ap) + v);
            V const mv = reinterpret_cast<const V*>(minv)[v];
            xv += alpha * pv;
            rv = rv - alpha * av;
            V const zv = mv * rv;
            static_for<VW>([&](auto jc) {
                constexpr int jj = decltype(jc)::value;
                rr += double(rv[jj]) * double(rv[jj]);
                rz += double(rv[jj]) * double(zv[jj]);
            });
            zk[k] = zv;
            if constexpr (KEEP_P) pk[k] = pv;
            store_stream<kVectorNT>(reinterpret_cast<V*>(x) + v, xv);
            reinterpret_cast<V*>(r)[v] = rv;
        }
        // one vector group's loads at a time: hoisting every group's loads ahead (the
        // scheduler's choice) needs more registers than the co-residency budget and spills
        __builtin_amdgcn_sched_barrier(0);
    });
    if (tail)
        for (int64_t i = full * VW; i < n; ++i)
        {
            x[i] += alpha * p[i];
            T const ri = r[i] - alpha * Ap[i];
            r[i] = ri;
            T const zi = minv[i] * ri;
            rr += double(ri) * double(ri);
            rz += double(ri) * double(zi);
        }
    rr = block_sum<kBlock>(rr);
    rz = block_sum<kBlock>(rz);
    if (threadIdx.x == 0)
    {
        part_rr[blockIdx.x] = rr;
        part_rz[blockIdx.x] = rz;
    }
    grid_barrier(bar);
    rr = reduce_partials<kBlock>(part_rr, gridDim.x);
    rz = reduce_partials<kBlock>(part_rz, gridDim.x);
    int32_t it_new;
    int32_t done = KRYLOV_RUNNING;
    if (MIXED)
    {
        it_new = s.it + 1;
        if (it_new >= max_it) done = KRYLOV_LIMIT;
    }
    else if (criterion == NEON_STOP_RELATIVE)
    {
        bool const conv = rr < s.threshold; // Eigen breaks before counting the iteration
        it_new = conv ? s.it : s.it + 1;
        if (conv)
            done = KRYLOV_CONVERGED;
        else if (it_new >= max_it)
            done = KRYLOV_LIMIT;
    }
    else
    {
        it_new = s.it + 1; // the CUDA loop counts at the top
        if (it_new >= max_it)
            done = KRYLOV_LIMIT;
        else if (!(sqrt(rz) > tol))
            done = KRYLOV_CONVERGED;
    }
    if (!isfinite(rr) || !isfinite(rz)) done = KRYLOV_NONFINITE;
    if (__hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) done = KRYLOV_NONFINITE;
    double const beta64 = rz / s.rz;
    T const beta = T(beta64);
    if (done == KRYLOV_RUNNING)
    {
        static_for<K>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            int64_t const v = g + k * total;
            if (v < full)
            {
                V pv;
                if constexpr (KEEP_P)
                    pv = pk[k];
                else
                    pv = reinterpret_cast<const V*>(p)[v];
                reinterpret_cast<V*>(p)[v] = zk[k] + beta * pv;
            }
        });
        if (tail)
            for (int64_t i = full * VW; i < n; ++i) p[i] = minv[i] * r[i] + beta * p[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        s.rr = rr;
        s.rz = rz;
        s.beta = beta64;
        s.it = it_new;
        s.done = done;
        if (MIXED) s.refine = 0;
        *st_next = s;
    }
}

// ------------------------------------------------------- generic vector kernels
__global__ __launch_bounds__(kBlock) void dot_kernel(int64_t n, const double* __restrict__ a,
                                                     const double* __restrict__ b, double* __restrict__ part)
{
    int64_t beg, end;
    xcd_chunk(n, beg, end);
    double s = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock) s += a[i] * b[i];
    s = block_sum<kBlock>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void reduce_to_kernel(const double* __restrict__ part, int npart,
                                                           double* __restrict__ out)
{
    double const s = reduce_partials<kBlock>(part, npart);
    if (threadIdx.x == 0) *out = s;
}

// y = a x + b y
__global__ __launch_bounds__(kBlock) void axpby_kernel(int64_t n, double a, const double* __restrict__ x, double b,
                                                       double* __restrict__ y)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}

// z = x - y
__global__ __launch_bounds__(kBlock) void sub_kernel(int64_t n, const double* __restrict__ x,
                                                     const double* __restrict__ y, double* __restrict__ z)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) z[i] = x[i] - y[i];
}

template <typename A, typename B>
__global__ __launch_bounds__(kBlock) void convert_kernel(int64_t n, const A* __restrict__ x, B* __restrict__ y)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) y[i] = B(x[i]);
}

// z = a x + b y
__global__ __launch_bounds__(kBlock) void lincomb_kernel(int64_t n, double a, const double* __restrict__ x, double b,
                                                         const double* __restrict__ y, double* __restrict__ z)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) z[i] = a * x[i] + b * y[i];
}

// z = d o x
__global__ __launch_bounds__(kBlock) void mul_kernel(int64_t n, const double* __restrict__ d,
                                                     const double* __restrict__ x, double* __restrict__ z)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) z[i] = d[i] * x[i];
}

// ------------------------------------------------ mixed precision CG (reliable updates)
// FP32 recurrence (x32, r32, p32, FP32 tangent values, FP32-accumulated SpMV) with FP64
// dot products, and FP64 reliable updates: when the recurrence residual has dropped by
// `delta` since the last update (or meets the threshold), x64 += x32, r = b - K x64 is
// recomputed in FP64 arithmetic and replaces r32; the search direction is kept.
// Convergence is decided only on the true FP64 residual.
__global__ __launch_bounds__(kBlock) void mixed_init_kernel(int64_t n, const double* __restrict__ b,
                                                            double* __restrict__ x64, float* __restrict__ r32,
                                                            float* __restrict__ x32, float* __restrict__ p32,
                                                            const float* __restrict__ minv,
                                                            double* __restrict__ part_rr,
                                                            double* __restrict__ part_rz)
{
    int64_t beg, end;
    xcd_chunk(n, beg, end);
    double rr = 0.0, rz = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock)
    {
        double const bi = b[i];
        float const ri = float(bi);
        float const zi = minv[i] * ri;
        x64[i] = 0.0;
        x32[i] = 0.0f;
        r32[i] = ri;
        p32[i] = zi;
        rr += bi * bi;
        rz += double(ri) * double(zi);
    }
    rr = block_sum<kBlock>(rr);
    rz = block_sum<kBlock>(rz);
    if (threadIdx.x == 0)
    {
        part_rr[blockIdx.x] = rr;
        part_rz[blockIdx.x] = rz;
    }
}

__global__ __launch_bounds__(kBlock) void mixed_init_state_kernel(const double* __restrict__ part_rr,
                                                                  const double* __restrict__ part_rz, int npart,
                                                                  krylov_state* __restrict__ st, double tol,
                                                                  int max_it)
{
    double const rr = reduce_partials<kBlock>(part_rr, npart);
    double const rz = reduce_partials<kBlock>(part_rz, npart);
    if (threadIdx.x != 0) return;
    krylov_state s{};
    s.rz = rz;
    s.rr = rr;
    s.rhs2 = rr;
    s.threshold = fmax(tol * tol * rr, DBL_MIN);
    s.aux0 = sqrt(rr);
    s.done = (rr == 0.0 || rr < s.threshold) ? KRYLOV_CONVERGED : KRYLOV_RUNNING;
    if (s.done == KRYLOV_RUNNING && max_it <= 0) s.done = KRYLOV_LIMIT;
    if (!isfinite(rr) || !isfinite(rz)) s.done = KRYLOV_NONFINITE;
    *st = s;
}

__global__ __launch_bounds__(kBlock) void mixed_update_xr_kernel(int64_t n, float* __restrict__ x32,
                                                                 float* __restrict__ r32,
                                                                 const float* __restrict__ p32,
                                                                 const float* __restrict__ Ap32,
                                                                 const float* __restrict__ minv,
                                                                 const double* __restrict__ part_pAp, int npart,
                                                                 const krylov_state* __restrict__ st,
                                                                 double* __restrict__ part_rr,
                                                                 double* __restrict__ part_rz)
{
    if (st->done) return;
    double const pAp = reduce_partials<kBlock>(part_pAp, npart);
    float const alpha = float(st->rz / pAp);
    int64_t beg, end;
    xcd_chunk_aligned<4>(n, beg, end);
    double rr = 0.0, rz = 0.0;
    auto body = [&](float& xi, float& ri, float pi, float api, float mi) {
        xi += alpha * pi;
        ri = ri - alpha * api;
        float const zi = mi * ri;
        rr += double(ri) * double(ri);
        rz += double(ri) * double(zi);
    };
    // four entries per thread in 16-byte accesses, then the chunk's tail
    int64_t const full = beg + (end - beg) / 4 * 4;
    for (int64_t i = beg + 4 * threadIdx.x; i < full; i += 4 * kBlock)
    {
        float4v xv = load_stream<kVectorNT>(reinterpret_cast<const float4v*>(x32 + i)),
                rv = *reinterpret_cast<const float4v*>(r32 + i);
        float4v const pv = *reinterpret_cast<const float4v*>(p32 + i),
                      av = load_stream<kVectorNT>(reinterpret_cast<const float4v*>(Ap32 + i)),
                      mv = *reinterpret_cast<const float4v*>(minv + i);
        float x0 = xv.x, x1 = xv.y, x2 = xv.z, x3 = xv.w, r0 = rv.x, r1 = rv.y, r2 = rv.z, r3 = rv.w;
        body(x0, r0, pv.x, av.x, mv.x);
        body(x1, r1, pv.y, av.y, mv.y);
        body(x2, r2, pv.z, av.z, mv.z);
        body(x3, r3, pv.w, av.w, mv.w);
        store_stream<kVectorNT>(reinterpret_cast<float4v*>(x32 + i), float4v{x0, x1, x2, x3});
        *reinterpret_cast<float4v*>(r32 + i) = float4v{r0, r1, r2, r3};
    }
    for (int64_t i = full + threadIdx.x; i < end; i += kBlock)
    {
        float xi = x32[i], ri = r32[i];
        body(xi, ri, p32[i], Ap32[i], minv[i]);
        x32[i] = xi;
        r32[i] = ri;
    }
    rr = block_sum<kBlock>(rr);
    rz = block_sum<kBlock>(rz);
    if (threadIdx.x == 0)
    {
        part_rr[blockIdx.x] = rr;
        part_rz[blockIdx.x] = rz;
    }
}

// single block: decide whether this iteration ends with a reliable update
__global__ __launch_bounds__(kBlock) void mixed_decide_kernel(const double* __restrict__ part_rr,
                                                              const double* __restrict__ part_rz, int npart,
                                                              krylov_state* __restrict__ st, double delta)
{
    if (st->done) return;
    double const rr = reduce_partials<kBlock>(part_rr, npart);
    double const rz = reduce_partials<kBlock>(part_rz, npart);
    if (threadIdx.x != 0) return;
    bool const refine = rr < st->threshold || sqrt(rr) < delta * st->aux0 || !isfinite(rr) || !isfinite(rz);
    st->refine = refine ? 1 : 0;
    st->rr = rr;
    st->aux1 = rz;
}

// x64 += x32, x32 = 0 (when a reliable update is due, or unconditionally with force)
__global__ __launch_bounds__(kBlock) void mixed_fold_kernel(int64_t n, double* __restrict__ x64,
                                                            float* __restrict__ x32,
                                                            const krylov_state* __restrict__ st, int force)
{
    if (!force && (st->done || !st->refine)) return;
    int64_t beg, end;
    xcd_chunk(n, beg, end);
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock)
    {
        x64[i] += double(x32[i]);
        x32[i] = 0.0f;
    }
}

// r32 = float(r64) after the FP64 residual; partials of r64.r64 and r32.z32
__global__ __launch_bounds__(kBlock) void mixed_refresh_kernel(int64_t n, const double* __restrict__ r64,
                                                               float* __restrict__ r32,
                                                               const float* __restrict__ minv,
                                                               const krylov_state* __restrict__ st,
                                                               double* __restrict__ part_rr,
                                                               double* __restrict__ part_rz)
{
    if (st->done || !st->refine) return;
    int64_t beg, end;
    xcd_chunk(n, beg, end);
    double rr = 0.0, rz = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock)
    {
        double const r = r64[i];
        float const ri = float(r);
        r32[i] = ri;
        rr += r * r;
        rz += double(ri) * double(minv[i] * ri);
    }
    rr = block_sum<kBlock>(rr);
    rz = block_sum<kBlock>(rz);
    if (threadIdx.x == 0)
    {
        part_rr[blockIdx.x] = rr;
        part_rz[blockIdx.x] = rz;
    }
}

// convergence on the true residual, beta, p32 = z + beta p32; block 0 writes the next state
// (decided = 0: no decision kernel ran this iteration, so r.r and r.z come from the partials
// of the x/r update and no reliable update is due)
__global__ __launch_bounds__(kBlock) void mixed_update_p_kernel(int64_t n, float* __restrict__ p32,
                                                                const float* __restrict__ r32,
                                                                const float* __restrict__ minv,
                                                                const double* __restrict__ part_rr_true,
                                                                const double* __restrict__ part_rz_true,
                                                                const double* __restrict__ part_rr,
                                                                const double* __restrict__ part_rz, int npart,
                                                                int decided, const krylov_state* __restrict__ st,
                                                                krylov_state* __restrict__ st_next, int max_it)
{
    krylov_state s = *st;
    if (s.done)
    {
        if (blockIdx.x == 0 && threadIdx.x == 0) *st_next = s;
        return;
    }
    double rr = s.rr, rz = s.aux1;
    if (!decided)
    {
        rr = reduce_partials<kBlock>(part_rr, npart);
        rz = reduce_partials<kBlock>(part_rz, npart);
        s.refine = 0;
    }
    bool conv = false;
    if (s.refine)
    {
        rr = reduce_partials<kBlock>(part_rr_true, npart);
        rz = reduce_partials<kBlock>(part_rz_true, npart);
        conv = rr < s.threshold;
        s.aux0 = sqrt(rr);
        s.restarts += 1; // counts reliable updates
    }
    int32_t const it_new = conv ? s.it : s.it + 1;
    int32_t done = conv ? KRYLOV_CONVERGED : (it_new >= max_it ? KRYLOV_LIMIT : KRYLOV_RUNNING);
    if (!isfinite(rr) || !isfinite(rz)) done = KRYLOV_NONFINITE;
    float const beta = float(rz / s.rz);
    if (done == KRYLOV_RUNNING)
    {
        int64_t beg, end;
        xcd_chunk_aligned<4>(n, beg, end);
        int64_t const full = beg + (end - beg) / 4 * 4;
        for (int64_t i = beg + 4 * threadIdx.x; i < full; i += 4 * kBlock)
        {
            float4v const mv = *reinterpret_cast<const float4v*>(minv + i),
                          rv = *reinterpret_cast<const float4v*>(r32 + i),
                          pv = *reinterpret_cast<const float4v*>(p32 + i);
            *reinterpret_cast<float4v*>(p32 + i) = float4v{mv.x * rv.x + beta * pv.x, mv.y * rv.y + beta * pv.y,
                                                           mv.z * rv.z + beta * pv.z, mv.w * rv.w + beta * pv.w};
        }
        for (int64_t i = full + threadIdx.x; i < end; i += kBlock) p32[i] = minv[i] * r32[i] + beta * p32[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        s.rr = rr;
        s.rz = rz;
        s.it = it_new;
        s.done = done;
        s.refine = 0;
        *st_next = s;
    }
}
} // namespace neon_amd
