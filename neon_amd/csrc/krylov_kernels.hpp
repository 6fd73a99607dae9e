// Device kernels of the Jacobi-preconditioned Krylov solvers: SpMV in three layouts (generic
// CSR, node-block CSR, the sliced node-block copy), vector updates fused with their dot-product
// partials, and the device-resident scalar state.
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

// y = A x (node-block layout, CSR-order values); with DOT, also writes the block partial of
// x . y.  MODE 0: y = Ax.  MODE 1: y = b - Ax (residual; b given in `bvec`).  Each XCD
// (workgroups b with b % 8 equal) owns one contiguous eighth of the nodes and every workgroup
// walks its own contiguous chunk of that eighth.  The operator for meshes whose nodes have too
// many neighbours for the sliced copy, and for the consistent mass.
template <int LPR, int ROWS, bool NT, typename VT, typename XT, typename YT, typename AT, bool DOT, int MODE>
__global__ __launch_bounds__(kBlock) void spmv_nb_kernel(nb_matrix A, const VT* __restrict__ val,
                                                         const XT* __restrict__ x, YT* __restrict__ y,
                                                         const double* __restrict__ bvec,
                                                         double* __restrict__ part,
                                                         const krylov_state* __restrict__ st, int gate)
{
    if (st && (st->done || (gate == 1 && !st->refine) || (gate == 2 && st->refine))) return;
    constexpr int G = kBlock / LPR; // lane groups per block
    int const g = threadIdx.x / LPR, l = threadIdx.x % LPR;
    int64_t beg, end;
    xcd_chunk(A.nodes, beg, end);
    int64_t const stride = G;
    int64_t const start = beg + g;
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
// reads one contiguous span.  Slice s holds the 64 consecutive nodes 64 s ... 64 s + 63 (lane l:
// node snode[64 s + l]; an empty lane past the last node, snode -1 - m, gathers like node m and
// holds zero values) and sptr[s+1] - sptr[s] steps, the largest block count among them; step kk
// of a node is its kk-th block (ascending column node, as in bcol), and steps past a node's
// block count hold zero values.  The 9 values of a step form one record of 9 x 64 values:
// value t = 3 i + j of lane l sits at
//   sval[576 (sptr[s] + kk) + sell_pos<VT>(t, l)]:
// the first 8 in groups of W = 16 / sizeof(VT) per lane (one 16-byte load per lane and group;
// group g spans 64 W values), the ninth after them (one value per lane).
// A step's column offsets (column node - lane node): scode[step] < 0 encodes one offset shared
// by the slice's lanes (shared_offset()) except at most two exception lanes, named by
// slane[s] (8 bits each, 0xff for none), whose 16-bit offsets are in sexc[step]; scode[step]
// >= 0 indexes 64 per-lane offsets at soff16 / soff32 [scode * 64 + lane].  A slice's steps
// are all of one kind (the product tests the first code of each chunk of 64 steps).
template <typename VT>
__host__ __device__ constexpr int sell_pos(int t, int lane)
{
    constexpr int W = 16 / int(sizeof(VT));
    return t < 8 ? (t / W) * 64 * W + lane * W + t % W : 512 + lane;
}

struct sell_matrix
{
    int32_t nodes;
    int32_t slices;
    const int32_t* sptr;   // [slices + 1] first step of each slice
    const int32_t* snode;  // [slices * 64] node of each lane (-1 - m: empty lane gathering like node m)
    const int32_t* slane;  // [slices] exception lanes of a slice's shared steps (8 bits each, 0xff: none)
    const int32_t* scode;  // [steps] shared_code(offset) or the step's index into the per-lane offsets
    const int32_t* sexc;   // [steps] 16-bit offsets of the exception lanes (shared steps)
    const int16_t* soff16; // per-lane offsets when every one fits 16 bits (else null)
    const int32_t* soff32; // per-lane offsets in 32 bits (used when soff16 is null)
};
constexpr int64_t kSharedBias = int64_t(1) << 30;
__host__ __device__ inline int32_t shared_code(int64_t offset)
{
    return int32_t(int64_t(INT32_MIN) + kSharedBias + offset);
}
__host__ __device__ inline int64_t shared_offset(int32_t code) { return int64_t(code) - INT32_MIN - kSharedBias; }

// y = A x on the sliced layout, one wavefront per slice, one node (three rows) per lane.  Each
// lane accumulates its three row sums over its steps in order, so no cross-lane reduction is
// needed.  With DOT, also writes the block partial of x . y (the stored, rounded y).
// MODE 0: y = Ax.  MODE 1: y = b - Ax.  Wavefront w of workgroup b takes slices 4 b + w, stride
// 4 gridDim (consecutive slices on different XCDs).  The slice index, its step range and each
// step's code are wavefront-uniform (scalar loads).
// NT: the values are read with nontemporal loads (streamed past the caches, which keep x).
template <bool I16, bool NT, typename VT, typename XT, typename YT, typename AT, bool DOT, int MODE>
__global__ __launch_bounds__(256) void spmv_sell_kernel(sell_matrix A, const VT* __restrict__ val,
                                                        const XT* __restrict__ x, YT* __restrict__ y,
                                                        const double* __restrict__ bvec, double* __restrict__ part,
                                                        const krylov_state* __restrict__ st, int gate)
{
    if (st && (st->done || (gate == 1 && !st->refine) || (gate == 2 && st->refine))) return;
    int const lane = threadIdx.x & 63;
    int32_t const wave = __builtin_amdgcn_readfirstlane(int32_t(threadIdx.x >> 6));
    double dot = 0.0;
    for (int32_t s = int32_t(blockIdx.x) * 4 + wave; s < A.slices; s += int32_t(gridDim.x) * 4)
    {
        int32_t const node = A.snode[int64_t(s) * 64 + lane];
        bool const ok = node >= 0;
        int64_t const nn = ok ? int64_t(node) : -1 - int64_t(node);
        int64_t const row = 3 * nn;
        int32_t const exl = A.slane[s];
        int const ex0 = exl & 0xff, ex1 = (exl >> 8) & 0xff;
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
        // the codes and exception words of up to 64 steps at a time, lane j holding step
        // kk + j's (coalesced loads); each step reads its own from that lane
        // one step's operands: its 9 values and the 3 entries of x at its column node
        struct step_t
        {
            VT v[9];
            XT x0, x1, x2;
        };
        auto fetch = [&](int64_t blk, int64_t m) {
            // the step's record: 8 values in 16-byte loads, the ninth alone
            constexpr int W = 16 / int(sizeof(VT));
            typedef VT vec_t __attribute__((ext_vector_type(W)));
            const VT* rec = val + blk * 576;
            step_t o;
#pragma unroll
            for (int g = 0; g < 8 / W; ++g)
            {
                const vec_t* src = reinterpret_cast<const vec_t*>(rec + g * 64 * W) + lane;
                vec_t q;
                if constexpr (NT)
                    q = __builtin_nontemporal_load(src);
                else
                    q = *src;
#pragma unroll
                for (int j = 0; j < W; ++j) o.v[g * W + j] = q[j];
            }
            if constexpr (NT)
                o.v[8] = __builtin_nontemporal_load(rec + 512 + lane);
            else
                o.v[8] = rec[512 + lane];
            o.x0 = x[3 * m];
            o.x1 = x[3 * m + 1];
            o.x2 = x[3 * m + 2];
            return o;
        };
        auto accumulate = [&](step_t const& o) {
            AT const x0 = AT(o.x0), x1 = AT(o.x1), x2 = AT(o.x2);
            AT const n0 = a0 + (AT(o.v[0]) * x0 + AT(o.v[1]) * x1 + AT(o.v[2]) * x2);
            AT const n1 = a1 + (AT(o.v[3]) * x0 + AT(o.v[4]) * x1 + AT(o.v[5]) * x2);
            AT const n2 = a2 + (AT(o.v[6]) * x0 + AT(o.v[7]) * x1 + AT(o.v[8]) * x2);
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
        };
        // steps in chunks of up to 64, whose codes and exception words one coalesced load each
        // brings in (lane j holds step c0 + j's); every step reads its own from that lane
        for (int32_t c0 = 0; c0 < ks; c0 += 64)
        {
            int32_t const cn = ks - c0 < 64 ? ks - c0 : 64;
            int64_t const base = int64_t(k0) + c0;
            bool const in = lane < cn;
            int32_t const codes = in ? A.scode[base + lane] : 0;
            int32_t const excs = in ? A.sexc[base + lane] : 0;
            if (__builtin_amdgcn_readlane(codes, 0) < 0)
            {
                // a slice of shared steps (all its steps are): the column offset comes from
                // registers, so later steps' loads are issued before step j is accumulated
                // (rotating register sets, no copies between them, which would wait for the
                // loads)
                auto shared = [&](int32_t j) {
                    int32_t const e = __builtin_amdgcn_readlane(excs, j);
                    int64_t const off = shared_offset(__builtin_amdgcn_readlane(codes, j));
                    int64_t const o0 = int16_t(e & 0xffff), o1 = int16_t(uint32_t(e) >> 16);
                    int64_t const o = lane == ex0 ? o0 : (lane == ex1 ? o1 : off);
                    return fetch(base + j, nn + o);
                };
                int32_t j = 0;
                if constexpr (sizeof(VT) == 8)
                {
                    // FP64: three register sets, so that while step j is accumulated steps
                    // j + 1 and j + 2 are in flight (at 170 VGPRs, two wavefronts per SIMD;
                    // 2.6 % faster than one step ahead at 128^3)
                    if (cn >= 5)
                    {
                        step_t sa = shared(0), sb = shared(1);
                        for (; j + 4 < cn; j += 3)
                        {
                            step_t const sc = shared(j + 2);
                            accumulate(sa);
                            sa = shared(j + 3);
                            accumulate(sb);
                            sb = shared(j + 4);
                            accumulate(sc);
                        }
                        accumulate(sa);
                        accumulate(sb);
                        j += 2;
                    }
                }
                else
                {
                    // FP32: two register sets, one step ahead (86 VGPRs; two steps ahead was
                    // 1.2 % slower at 128^3)
                    if (cn >= 2)
                    {
                        step_t sa = shared(0);
                        for (; j + 2 < cn; j += 2)
                        {
                            step_t const sb = shared(j + 1);
                            accumulate(sa);
                            sa = shared(j + 2);
                            accumulate(sb);
                        }
                        accumulate(sa);
                        j += 1;
                    }
                }
                for (; j < cn; ++j) accumulate(shared(j));
            }
            else
            {
                // per-lane offsets (each step's gathers wait for its offset load)
                for (int32_t j = 0; j < cn; ++j)
                {
                    int64_t const code = __builtin_amdgcn_readlane(codes, j);
                    int64_t m;
                    if constexpr (I16)
                        m = nn + int64_t(A.soff16[code * 64 + lane]);
                    else
                        m = nn + int64_t(A.soff32[code * 64 + lane]);
                    accumulate(fetch(base + j, m));
                }
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
        dot = block_sum<256>(dot);
        if (threadIdx.x == 0) part[blockIdx.x] = dot;
    }
}

// Sliced copy of node-block values (layout above), with coalesced reads: one 256-thread
// workgroup per slice; for each dof row i the slice's 64 row segments (3 deg contiguous values
// each, CSR order val[9 bptr[n] + i (3 deg) + 3 kk + j]) are read into LDS (rows of `row_len`
// = 3 * max degree values), then written to the slice's step records (sell_pos).  Steps past a
// node's block count, and empty lanes, get zeros.  Dynamic LDS: 64 * row_len values.
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
        int32_t const n = A.snode[int64_t(s) * 64 + tid];
        bool const ok = n >= 0;
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
            sval[(int64_t(k0) + kk) * 576 + sell_pos<VT>(3 * i + j, lane)] = v;
        }
        __syncthreads();
    }
}

// ------------------------------------------------ symmetric sliced layout
// For a symmetric tangent the copy keeps the diagonal block and the upper blocks (column node -
// row node = o > 0) only.  Slice s holds nodes 64 s .. 64 s + 63 (one per lane); its records of
// 9 x 64 values (sell_pos layout) are the diagonal block (record rdiag[s]) and one record per
// offset o in the slice's direct list (the upper offsets of its lanes' nodes): lane l's block
// (i, i + o), zero where node i has no such neighbour.  The product adds, for every lane node i,
// K(i, i + o) x(i + o) for the direct offsets and the transpose of K(i - o, i) x(i - o) for the
// slice's lower offsets (those of its lanes' lower neighbours): that block is lane (i - o) % 64
// of the record for o of slice (i - o) / 64, which is one of the two slices floor((64 s - o) /
// 64) and the next one, whose record indices the entry carries.  Every off-diagonal value is
// read twice, and the schedule makes the second read follow the first closely enough to be
// served by the L2.  A slice's entries merge the two lists by offset, each {o, direct record or
// -1, record in slice floor((64 s - o) / 64) or -1, record in the next slice or -1}, so that the
// direct and the transposed read of one offset are issued together.
// Schedule: wave slot q (workgroup q / 4, on XCD (q / 4) % 8) takes slice sched[q] (-1: none).
// The host orders the slices in tiles (a band of the in-plane index, then plane by plane) and
// gives each XCD one contiguous eighth of that order, so a slice's neighbours one plane down
// were read on the same XCD a few workgroups earlier.
constexpr int kMaxSymEntries = 40; // <= 64: a slice's entries fit one wavefront's lanes
typedef int int4v __attribute__((ext_vector_type(4)));
struct sym_matrix
{
    int32_t nodes;
    int32_t slices;
    int32_t slots;          // wave slots of the schedule (a multiple of 32)
    int64_t records;        // records of the copy
    const int32_t* sched;   // [slots] slice of each wave slot, -1 for none
    const int32_t* rdiag;   // [slices] record of each slice's diagonal block
    const int32_t* eptr;    // [slices + 1] entry range of each slice
    const int4v* ent;       // entries, by slice and ascending offset
    const int32_t* rslice;  // [records] slice of each record (packing)
    const int32_t* roff;    // [records] offset of each record, 0 for the diagonal (packing)
};

// y = A x on the symmetric sliced layout (DOT, MODE, gate as in spmv_sell_kernel).  FP32
// accumulation rotates over four accumulator sets as there; FP64 uses one.
template <typename VT, typename XT, typename YT, typename AT, bool DOT, int MODE>
__global__ __launch_bounds__(256) void spmv_sym_kernel(sym_matrix A, const VT* __restrict__ val,
                                                       const XT* __restrict__ x, YT* __restrict__ y,
                                                       const double* __restrict__ bvec, double* __restrict__ part,
                                                       const krylov_state* __restrict__ st, int gate)
{
    if (st && (st->done || (gate == 1 && !st->refine) || (gate == 2 && st->refine))) return;
    int const lane = threadIdx.x & 63;
    int32_t const wave = __builtin_amdgcn_readfirstlane(int32_t(threadIdx.x >> 6));
    int64_t const nodes = A.nodes;
    constexpr int W = 16 / int(sizeof(VT));
    typedef VT vec_t __attribute__((ext_vector_type(W)));
    // the 9 values of lane l of the record at rec (NT: nontemporal, for a record read once)
    auto load_rec = [&](const VT* rec, int l, VT(&v)[9], auto nt) {
#pragma unroll
        for (int g = 0; g < 8 / W; ++g)
        {
            const vec_t* src = reinterpret_cast<const vec_t*>(rec + g * 64 * W) + l;
            vec_t q;
            if constexpr (decltype(nt)::value)
                q = __builtin_nontemporal_load(src);
            else
                q = *src;
#pragma unroll
            for (int j = 0; j < W; ++j) v[g * W + j] = q[j];
        }
        if constexpr (decltype(nt)::value)
            v[8] = __builtin_nontemporal_load(rec + 512 + l);
        else
            v[8] = rec[512 + l];
    };
    double dot = 0.0;
    for (int32_t q = int32_t(blockIdx.x) * 4 + wave; q < A.slots; q += int32_t(gridDim.x) * 4)
    {
        int32_t const s = __builtin_amdgcn_readfirstlane(A.sched[q]);
        if (s < 0) continue;
        int64_t const i = int64_t(s) * 64 + lane;
        bool const ok = i < nodes;
        int64_t const ic = ok ? i : nodes - 1;
        int64_t const row = 3 * ic;
        int64_t const rdiag = A.rdiag[s];
        double br0 = 0.0, br1 = 0.0, br2 = 0.0;
        if constexpr (MODE == 1)
        {
            br0 = bvec[row];
            br1 = bvec[row + 1];
            br2 = bvec[row + 2];
        }
        XT const xo0 = x[row], xo1 = x[row + 1], xo2 = x[row + 2];
        AT a0 = AT(0), a1 = AT(0), a2 = AT(0), b0 = AT(0), b1 = AT(0), b2 = AT(0);
        AT c0 = AT(0), c1 = AT(0), c2 = AT(0), d0 = AT(0), d1 = AT(0), d2 = AT(0);
        // one block's product (transposed: K^T x) into the next accumulator set
        auto accumulate = [&](VT const(&v)[9], XT x0_, XT x1_, XT x2_, bool tr) {
            AT const x0 = AT(x0_), x1 = AT(x1_), x2 = AT(x2_);
            AT n0, n1, n2;
            if (tr)
            {
                n0 = a0 + (AT(v[0]) * x0 + AT(v[3]) * x1 + AT(v[6]) * x2);
                n1 = a1 + (AT(v[1]) * x0 + AT(v[4]) * x1 + AT(v[7]) * x2);
                n2 = a2 + (AT(v[2]) * x0 + AT(v[5]) * x1 + AT(v[8]) * x2);
            }
            else
            {
                n0 = a0 + (AT(v[0]) * x0 + AT(v[1]) * x1 + AT(v[2]) * x2);
                n1 = a1 + (AT(v[3]) * x0 + AT(v[4]) * x1 + AT(v[5]) * x2);
                n2 = a2 + (AT(v[6]) * x0 + AT(v[7]) * x1 + AT(v[8]) * x2);
            }
            if constexpr (sizeof(AT) == 4)
            {
                a0 = b0, a1 = b1, a2 = b2;
                b0 = c0, b1 = c1, b2 = c2;
                c0 = d0, c1 = d1, c2 = d2;
                d0 = n0, d1 = n1, d2 = n2;
            }
            else
            {
                a0 = n0, a1 = n1, a2 = n2;
            }
        };
        {
            VT v[9];
            load_rec(val + rdiag * 576, lane, v, std::true_type{});
            accumulate(v, xo0, xo1, xo2, false);
        }
        // one entry's operands: the stored block K(i, i + o) and x(i + o) (zero where the slice
        // has no direct record for o), and the transpose of K(i - o, i) and x(i - o) (x counts
        // as zero where node i - o does not exist or holds no block for o: its values are then
        // read from the slice's own diagonal record, in bounds)
        int32_t const e0 = A.eptr[s], ne = A.eptr[s + 1] - e0;
        // the slice's entries (at most kMaxSymEntries <= 64), lane k holding entry k: one
        // coalesced load, then each entry read from its lane (no scalar-load wait in the loop)
        int4v const mine = lane < ne ? A.ent[e0 + lane] : int4v{0, 0, -1, -1};
        struct pair_t
        {
            VT v[9], w[9];
            XT x0, x1, x2, z0, z1, z2;
            bool dv, jv;
        };
        auto fetch = [&](int32_t m) {
            pair_t o;
            int4v const e{__builtin_amdgcn_readlane(mine.x, m), __builtin_amdgcn_readlane(mine.y, m),
                          __builtin_amdgcn_readlane(mine.z, m), __builtin_amdgcn_readlane(mine.w, m)};
            int64_t const off = e.x;
            int64_t const rd = e.y;
            o.dv = rd >= 0;
            int64_t const mi = i + off < nodes ? i + off : nodes - 1;
            load_rec(val + (o.dv ? rd : rdiag) * 576, lane, o.v, std::false_type{});
            o.x0 = x[3 * mi], o.x1 = x[3 * mi + 1], o.x2 = x[3 * mi + 2];
            int64_t const j0 = i - off;
            int64_t const sa = (int64_t(s) * 64 - off) >> 6; // floor division by 64
            int64_t const rt = (j0 >> 6) == sa ? int64_t(e.z) : int64_t(e.w);
            o.jv = j0 >= 0 && j0 < nodes && rt >= 0;
            int64_t const j = o.jv ? j0 : ic;
            load_rec(val + (o.jv ? rt : rdiag) * 576, o.jv ? int(j & 63) : lane, o.w, std::false_type{});
            o.z0 = x[3 * j], o.z1 = x[3 * j + 1], o.z2 = x[3 * j + 2];
            return o;
        };
        auto accumulate_pair = [&](pair_t const& o) {
            accumulate(o.v, o.dv ? o.x0 : XT(0), o.dv ? o.x1 : XT(0), o.dv ? o.x2 : XT(0), false);
            accumulate(o.w, o.jv ? o.z0 : XT(0), o.jv ? o.z1 : XT(0), o.jv ? o.z2 : XT(0), true);
        };
        // software-pipelined over two register sets: entry m + 1's loads are issued before
        // entry m is accumulated (3 % faster at 128^3 in both precisions than a plain loop)
        int32_t m = 0;
        if (ne >= 2)
        {
            pair_t sa_ = fetch(0);
            for (; m + 2 < ne; m += 2)
            {
                pair_t const sb = fetch(m + 1);
                accumulate_pair(sa_);
                sa_ = fetch(m + 2);
                accumulate_pair(sb);
            }
            accumulate_pair(sa_);
            m += 1;
        }
        for (; m < ne; ++m) accumulate_pair(fetch(m));
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
            dot += ok ? double(xo0) * double(y0) + double(xo1) * double(y1) + double(xo2) * double(y2) : 0.0;
    }
    if constexpr (DOT)
    {
        dot = block_sum<256>(dot);
        if (threadIdx.x == 0) part[blockIdx.x] = dot;
    }
}

// Symmetric sliced copy of node-block values (CSR order val[9 bptr[n] + i (3 deg) + 3 kk + j]):
// one wavefront per record; each lane finds its column node (itself for a diagonal record, else
// i + the record's offset) among its row's blocks, zeros where the row has no such block.  With
// minv (a positive Jacobi inverse) the copy holds the symmetrically scaled tangent
// K~(a, b) = sqrt(minv[a]) K(a, b) sqrt(minv[b]) (scaled in FP64, then rounded to VT).
template <typename VT>
__global__ __launch_bounds__(256) void sym_pack_kernel(sym_matrix A, const int32_t* __restrict__ bptr,
                                                       const int32_t* __restrict__ bcol, const VT* __restrict__ val,
                                                       VT* __restrict__ sval, const double* __restrict__ minv)
{
    int64_t const rec = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (rec >= A.records) return;
    int const lane = threadIdx.x & 63;
    int64_t const s = A.rslice[rec];
    int64_t const i = s * 64 + lane;
    VT v[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) v[t] = VT(0);
    if (i < A.nodes)
    {
        int64_t const m = i + A.roff[rec];
        int32_t const b0 = bptr[i], deg = bptr[i + 1] - b0;
        for (int32_t kk = 0; kk < deg; ++kk)
            if (bcol[b0 + kk] == m)
            {
                const VT* src = val + 9ll * b0 + 3 * kk;
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) v[3 * r + c] = src[int64_t(r) * 3 * deg + c];
                if (minv)
                {
#pragma unroll
                    for (int r = 0; r < 3; ++r)
                    {
                        double const sr = sqrt(minv[3 * i + r]);
#pragma unroll
                        for (int c = 0; c < 3; ++c)
                            v[3 * r + c] = VT(double(v[3 * r + c]) * sr * sqrt(minv[3 * m + c]));
                    }
                }
                break;
            }
    }
    VT* dst = sval + rec * 576;
#pragma unroll
    for (int t = 0; t < 9; ++t) dst[sell_pos<VT>(t, lane)] = v[t];
}

// Generic CSR (any sparsity): LPR lanes per row.
template <int LPR, typename VT, typename XT, typename YT, typename AT, bool DOT, int MODE>
__global__ __launch_bounds__(kBlock) void spmv_csr_kernel(csr_matrix A, const VT* __restrict__ val,
                                                          const XT* __restrict__ x, YT* __restrict__ y,
                                                          const double* __restrict__ bvec,
                                                          double* __restrict__ part,
                                                          const krylov_state* __restrict__ st, int gate)
{
    if (st && (st->done || (gate == 1 && !st->refine) || (gate == 2 && st->refine))) return;
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

// part_a[kMaxPartials] = the sum of part_a[0, n) (and the same for part_b when given), in one
// 1024-thread block, in a fixed order; the consumers then read one total (partials_total with
// n < 0) instead of every block reducing all n partials.  Skipped when the state is done.
__global__ __launch_bounds__(1024) void totals_kernel(double* __restrict__ part_a, double* __restrict__ part_b,
                                                       int n, const krylov_state* __restrict__ st)
{
    if (st && st->done) return;
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < n; k += 1024)
    {
        a += part_a[k];
        if (part_b) b += part_b[k];
    }
    a = block_sum<1024>(a);
    if (part_b) b = block_sum<1024>(b);
    if (threadIdx.x == 0)
    {
        part_a[kMaxPartials] = a;
        if (part_b) part_b[kMaxPartials] = b;
    }
}

// r -= alpha Ap with the partials of r.r and r.z (z = M^-1 r); x moves in the p update, which
// reads p anyway (one vector pass fewer).
__global__ __launch_bounds__(kBlock) void cg_update_xr_kernel(int64_t n, double* __restrict__ r,
                                                              const double* __restrict__ Ap,
                                                              const double* __restrict__ minv,
                                                              const double* __restrict__ part_pAp, int npart,
                                                              const krylov_state* __restrict__ st,
                                                              double* __restrict__ part_rr,
                                                              double* __restrict__ part_rz)
{
    if (st->done) return;
    double const pAp = partials_total<kBlock>(part_pAp, npart);
    double const alpha = st->rz / pAp;
    int64_t beg, end;
    xcd_chunk_aligned<2>(n, beg, end);
    double rr = 0.0, rz = 0.0;
    auto body = [&](double& ri, double api, double mi) {
        ri = ri - alpha * api;
        double const zi = mi * ri;
        rr += ri * ri;
        rz += ri * zi;
    };
    // two entries per thread in 16-byte accesses, then the chunk's odd tail
    int64_t const full = beg + (end - beg) / 2 * 2;
    for (int64_t i = beg + 2 * threadIdx.x; i < full; i += 2 * kBlock)
    {
        double2v rv = *reinterpret_cast<const double2v*>(r + i);
        double2v const av = load_stream<kVectorNT>(reinterpret_cast<const double2v*>(Ap + i)),
                       mv = *reinterpret_cast<const double2v*>(minv + i);
        double r0 = rv.x, r1 = rv.y;
        body(r0, av.x, mv.x);
        body(r1, av.y, mv.y);
        *reinterpret_cast<double2v*>(r + i) = double2v{r0, r1};
    }
    for (int64_t i = full + threadIdx.x; i < end; i += kBlock)
    {
        double ri = r[i];
        body(ri, Ap[i], minv[i]);
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

// x += alpha p (every iteration that ran: both stopping rules update x before their test), the
// convergence decision, beta, p = z + beta p; block 0 writes the next scalar state.
__global__ __launch_bounds__(kBlock) void cg_update_p_kernel(int64_t n, double* __restrict__ p,
                                                             double* __restrict__ x,
                                                             const double* __restrict__ r,
                                                             const double* __restrict__ minv,
                                                             const double* __restrict__ part_pAp,
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
    double const alpha = s.rz / partials_total<kBlock>(part_pAp, -1);
    double const rr = partials_total<kBlock>(part_rr, npart);
    double const rz = partials_total<kBlock>(part_rz, npart);
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
    bool const more = done == KRYLOV_RUNNING;
    int64_t beg, end;
    xcd_chunk_aligned<2>(n, beg, end);
    int64_t const full = beg + (end - beg) / 2 * 2;
    for (int64_t i = beg + 2 * threadIdx.x; i < full; i += 2 * kBlock)
    {
        double2v const pv = *reinterpret_cast<const double2v*>(p + i);
        double2v const xv = load_stream<kVectorNT>(reinterpret_cast<const double2v*>(x + i));
        store_stream<kVectorNT>(reinterpret_cast<double2v*>(x + i), double2v{xv.x + alpha * pv.x, xv.y + alpha * pv.y});
        if (more)
        {
            double2v const mv = *reinterpret_cast<const double2v*>(minv + i),
                           rv = *reinterpret_cast<const double2v*>(r + i);
            *reinterpret_cast<double2v*>(p + i) = double2v{mv.x * rv.x + beta * pv.x, mv.y * rv.y + beta * pv.y};
        }
    }
    for (int64_t i = full + threadIdx.x; i < end; i += kBlock)
    {
        double const pi = p[i];
        x[i] = x[i] + alpha * pi;
        if (more) p[i] = minv[i] * r[i] + beta * pi;
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


// ------------------------------------------------------------- BiCGSTAB (FP64)
// Device-resident right-preconditioned BiCGSTAB: every scalar stays in one krylov_state, which
// single-workgroup "decision" kernels update between the grid kernels that read it (kernel
// boundaries order the two), and the host polls it like the CG loop.  Two stopping rules:
//   CUDA (biconjugate_gradient_stabilised_cuda.cpp:62-191): the loop runs while
//     sqrt(r0 . r) > tol and k < max; it stops early when ||s|| < tol (x += alpha p_hat);
//     omega == 0 is a breakdown; k >= max after the loop fails even if the last test passed.
//   Eigen (BiCGSTAB.h as called by linear_solver.cpp:111-138): while |r|^2 > tol^2 |b|^2 and
//     i < max; rho = r0 . r, with a restart (r = b - A x, r0 = r, i = 0 the first time) when
//     |rho| < eps^2 |r0|^2; omega = t.s / t.t (0 when t.t = 0).
// State fields: rz = rho (CUDA's "residual" r0.r), aux0 = rho_old, alpha, omega, rr = |r|^2
// (Eigen), rhs2 = |b|^2, threshold = tol^2 |b|^2, aux1 = |s| (CUDA) or |r0|^2 (Eigen),
// refine = 1 while this iteration's early exit (CUDA) or restart (Eigen) is due, it = k / i.
enum : int
{
    BICG_GATE_NONE = 0,
    BICG_GATE_IF_FLAG = 1,     // run only when refine is set
    BICG_GATE_UNLESS_FLAG = 2, // run only when refine is clear
};
__device__ inline bool bicg_skip(const krylov_state* st, int gate)
{
    return st->done || (gate == BICG_GATE_IF_FLAG && !st->refine) || (gate == BICG_GATE_UNLESS_FLAG && st->refine);
}

// partials of a.b (and of c.d when c is set), skipped per gate
__global__ __launch_bounds__(kBlock) void bicg_dot_kernel(int64_t n, const double* __restrict__ a,
                                                          const double* __restrict__ b, const double* __restrict__ c,
                                                          const double* __restrict__ d, double* __restrict__ part_ab,
                                                          double* __restrict__ part_cd,
                                                          const krylov_state* __restrict__ st, int gate)
{
    if (bicg_skip(st, gate)) return;
    int64_t beg, end;
    xcd_chunk(n, beg, end);
    double s = 0.0, t = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock)
    {
        s += a[i] * b[i];
        if (c) t += c[i] * d[i];
    }
    s = block_sum<kBlock>(s);
    if (c) t = block_sum<kBlock>(t);
    if (threadIdx.x == 0)
    {
        part_ab[blockIdx.x] = s;
        if (c) part_cd[blockIdx.x] = t;
    }
}

// p = r (first iteration of the CUDA rule) or p = r + beta (p - omega v), beta = (rho / rho_old)
// (alpha / omega); p_hat = M^-1 p
__global__ __launch_bounds__(kBlock) void bicg_p_kernel(int64_t n, double* __restrict__ p,
                                                        const double* __restrict__ r, const double* __restrict__ v,
                                                        const double* __restrict__ minv, double* __restrict__ p_hat,
                                                        const krylov_state* __restrict__ st, int first_plain)
{
    krylov_state const s = *st;
    if (s.done) return;
    bool const plain = first_plain && s.it == 0;
    double const beta = (s.rz / s.aux0) * (s.alpha / s.omega);
    int64_t beg, end;
    xcd_chunk(n, beg, end);
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock)
    {
        double const pi = plain ? r[i] : r[i] + beta * (p[i] - s.omega * v[i]);
        p[i] = pi;
        p_hat[i] = minv[i] * pi;
    }
}

// alpha = rho / (r0 . v); s = r - alpha v; s_hat = M^-1 s; partials of s.s
__global__ __launch_bounds__(kBlock) void bicg_s_kernel(int64_t n, const double* __restrict__ r,
                                                        const double* __restrict__ v, const double* __restrict__ minv,
                                                        double* __restrict__ sv, double* __restrict__ s_hat,
                                                        const double* __restrict__ part_r0v, int npart,
                                                        double* __restrict__ part_ss,
                                                        const krylov_state* __restrict__ st)
{
    if (st->done) return;
    double const alpha = st->rz / reduce_partials<kBlock>(part_r0v, npart);
    int64_t beg, end;
    xcd_chunk(n, beg, end);
    double ss = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock)
    {
        double const si = r[i] - alpha * v[i];
        sv[i] = si;
        s_hat[i] = minv[i] * si;
        ss += si * si;
    }
    ss = block_sum<kBlock>(ss);
    if (threadIdx.x == 0) part_ss[blockIdx.x] = ss;
}

// single block: alpha, |s| and the CUDA rule's early exit (refine = 1 when |s| < tol)
__global__ __launch_bounds__(kBlock) void bicg_alpha_kernel(const double* __restrict__ part_r0v,
                                                            const double* __restrict__ part_ss, int npart,
                                                            krylov_state* __restrict__ st, int cuda_rule, double tol)
{
    if (st->done) return;
    double const r0v = reduce_partials<kBlock>(part_r0v, npart);
    double const ss = reduce_partials<kBlock>(part_ss, npart);
    if (threadIdx.x != 0) return;
    st->alpha = st->rz / r0v;
    st->aux1 = sqrt(ss);
    st->refine = cuda_rule && sqrt(ss) < tol ? 1 : 0;
}

// x += alpha p_hat (+ omega s_hat); unless the early exit is due: omega from t.s and t.t,
// r = s - omega t, partials of r0.r and r.r
__global__ __launch_bounds__(kBlock) void bicg_update_kernel(int64_t n, double* __restrict__ x,
                                                             double* __restrict__ r, const double* __restrict__ r0,
                                                             const double* __restrict__ p_hat,
                                                             const double* __restrict__ s_hat,
                                                             const double* __restrict__ sv,
                                                             const double* __restrict__ t,
                                                             const double* __restrict__ part_ts,
                                                             const double* __restrict__ part_tt, int npart,
                                                             double* __restrict__ part_r0r,
                                                             double* __restrict__ part_rr,
                                                             const krylov_state* __restrict__ st, int cuda_rule)
{
    krylov_state const s = *st;
    if (s.done) return;
    int64_t beg, end;
    xcd_chunk(n, beg, end);
    if (s.refine)
    {
        for (int64_t i = beg + threadIdx.x; i < end; i += kBlock) x[i] += s.alpha * p_hat[i];
        return;
    }
    double const ts = reduce_partials<kBlock>(part_ts, npart), tt = reduce_partials<kBlock>(part_tt, npart);
    double const omega = cuda_rule ? ts / tt : (tt > 0.0 ? ts / tt : 0.0);
    double r0r = 0.0, rr = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock)
    {
        if (cuda_rule)
        {
            // x = x + alpha p_hat + omega s_hat as two updates (the two cublasDaxpy calls)
            double xi = x[i] + s.alpha * p_hat[i];
            x[i] = xi + omega * s_hat[i];
        }
        else
            x[i] += s.alpha * p_hat[i] + omega * s_hat[i]; // x += (alpha y + w z)
        double const ri = sv[i] - omega * t[i];
        r[i] = ri;
        r0r += r0[i] * ri;
        rr += ri * ri;
    }
    r0r = block_sum<kBlock>(r0r);
    rr = block_sum<kBlock>(rr);
    if (threadIdx.x == 0)
    {
        part_r0r[blockIdx.x] = r0r;
        part_rr[blockIdx.x] = rr;
    }
}

// single block: the end of an iteration (counts, scalars, the loop tests)
__global__ __launch_bounds__(kBlock) void bicg_finish_kernel(const double* __restrict__ part_ts,
                                                             const double* __restrict__ part_tt,
                                                             const double* __restrict__ part_r0r,
                                                             const double* __restrict__ part_rr, int npart,
                                                             krylov_state* __restrict__ st, int cuda_rule,
                                                             double tol, int max_it)
{
    if (st->done) return;
    double ts = 0.0, tt = 0.0, r0r = 0.0, rr = 0.0;
    if (!st->refine)
    {
        ts = reduce_partials<kBlock>(part_ts, npart);
        tt = reduce_partials<kBlock>(part_tt, npart);
        r0r = reduce_partials<kBlock>(part_r0r, npart);
        rr = reduce_partials<kBlock>(part_rr, npart);
    }
    if (threadIdx.x != 0) return;
    krylov_state s = *st;
    int32_t const k = s.it + 1;
    if (cuda_rule)
    {
        if (s.refine)
        {
            // ||s|| < tol: x += alpha p_hat, residual = ||s||^2, leave the loop
            s.rz = s.aux1 * s.aux1;
            s.done = k >= max_it ? KRYLOV_LIMIT : KRYLOV_CONVERGED;
        }
        else
        {
            s.omega = ts / tt;
            s.aux0 = s.rz;
            s.rz = r0r;
            if (s.omega == 0.0) // is_approx(omega, 0.0): an exact zero
                s.done = KRYLOV_BREAKDOWN;
            else if (!(sqrt(s.rz) > tol) || k >= max_it)
                s.done = k >= max_it ? KRYLOV_LIMIT : KRYLOV_CONVERGED;
            if (!isfinite(s.rz) && s.done != KRYLOV_BREAKDOWN) s.done = KRYLOV_NONFINITE;
        }
        s.rr = rr;
    }
    else
    {
        s.omega = tt > 0.0 ? ts / tt : 0.0;
        s.rr = rr;
        s.aux0 = s.rz; // rho_old for the next iteration
        s.rz = r0r;    // r0 . r, the next iteration's rho unless it restarts
        if (!(rr > s.threshold))
            s.done = KRYLOV_CONVERGED;
        else if (k >= max_it)
            s.done = KRYLOV_LIMIT;
        if (!isfinite(rr)) s.done = KRYLOV_NONFINITE;
    }
    s.it = k;
    s.refine = 0;
    *st = s;
}

// Eigen rule, single block at the top of an iteration: the restart test |rho| < eps^2 |r0|^2
// (refine = 1 when a restart is due)
__global__ __launch_bounds__(kBlock) void bicg_restart_test_kernel(krylov_state* __restrict__ st)
{
    if (threadIdx.x != 0 || st->done) return;
    double const eps2 = DBL_EPSILON * DBL_EPSILON;
    st->refine = fabs(st->rz) < eps2 * st->aux1 ? 1 : 0;
}

// Eigen rule, restart (gated on refine): r0 = r (the fresh residual), partials of r.r
__global__ __launch_bounds__(kBlock) void bicg_restart_kernel(int64_t n, const double* __restrict__ r,
                                                              double* __restrict__ r0, double* __restrict__ part_rr,
                                                              const krylov_state* __restrict__ st)
{
    if (bicg_skip(st, BICG_GATE_IF_FLAG)) return;
    int64_t beg, end;
    xcd_chunk(n, beg, end);
    double rr = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock)
    {
        r0[i] = r[i];
        rr += r[i] * r[i];
    }
    rr = block_sum<kBlock>(rr);
    if (threadIdx.x == 0) part_rr[blockIdx.x] = rr;
}

// Eigen rule, single block after a restart: rho = |r0|^2 = |r|^2, and the first restart resets
// the iteration count
__global__ __launch_bounds__(kBlock) void bicg_restart_finish_kernel(const double* __restrict__ part_rr, int npart,
                                                                     krylov_state* __restrict__ st)
{
    if (st->done || !st->refine) return;
    double const rr = reduce_partials<kBlock>(part_rr, npart);
    if (threadIdx.x != 0) return;
    st->rz = rr;
    st->aux1 = rr;
    if (st->restarts++ == 0) st->it = 0;
    st->refine = 0;
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

// The same start for the Jacobi-scaled system K~ x~ = b~ (K~ = D^-1/2 K D^-1/2, b~ = D^-1/2 b,
// x = D^-1/2 x~, D^-1/2 = sqrt(minv64)) that the FP32 recurrence runs on: r32 = p32 =
// float(b~); x64 (unscaled) = 0; the partials of b.b (the stopping rule's reference) and of
// r32.r32 (= r.z)
__global__ __launch_bounds__(kBlock) void mixed_init_scaled_kernel(int64_t n, const double* __restrict__ b,
                                                                   double* __restrict__ x64, float* __restrict__ r32,
                                                                   float* __restrict__ x32, float* __restrict__ p32,
                                                                   const double* __restrict__ minv64,
                                                                   double* __restrict__ part_rr,
                                                                   double* __restrict__ part_rz)
{
    int64_t beg, end;
    xcd_chunk(n, beg, end);
    double rr = 0.0, rz = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock)
    {
        double const bi = b[i];
        float const ri = float(sqrt(minv64[i]) * bi);
        x64[i] = 0.0;
        x32[i] = 0.0f;
        r32[i] = ri;
        p32[i] = ri;
        rr += bi * bi;
        rz += double(ri) * double(ri);
    }
    rr = block_sum<kBlock>(rr);
    rz = block_sum<kBlock>(rz);
    if (threadIdx.x == 0)
    {
        part_rr[blockIdx.x] = rr;
        part_rz[blockIdx.x] = rz;
    }
}

// block partials of the number of entries of v that are not positive and finite
__global__ __launch_bounds__(kBlock) void nonpositive_count_kernel(int64_t n, const double* __restrict__ v,
                                                                   double* __restrict__ part)
{
    int64_t beg, end;
    xcd_chunk(n, beg, end);
    double c = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock) c += (v[i] > 0.0 && isfinite(v[i])) ? 0.0 : 1.0;
    c = block_sum<kBlock>(c);
    if (threadIdx.x == 0) part[blockIdx.x] = c;
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

// r32 -= alpha Ap32 with FP64 partials of r.r and r.z; x32 moves in the p update (or in the
// fold of a reliable update).  SCALED: the recurrence of the Jacobi-scaled system
// (r~ = D^-1/2 r, z~ = r~, no minv read): the partials of r~.r~ (= r.z) only, and with RR also
// those of r.r = sum r~_i^2 / minv_i, which the reliable-update test reads (minv is read then)
template <bool SCALED, bool RR>
__global__ __launch_bounds__(kBlock) void mixed_update_xr_kernel(int64_t n, float* __restrict__ r32,
                                                                 const float* __restrict__ Ap32,
                                                                 const float* __restrict__ minv,
                                                                 const double* __restrict__ part_pAp, int npart,
                                                                 const krylov_state* __restrict__ st,
                                                                 double* __restrict__ part_rr,
                                                                 double* __restrict__ part_rz)
{
    if (st->done) return;
    double const pAp = partials_total<kBlock>(part_pAp, npart);
    float const alpha = float(st->rz / pAp);
    int64_t beg, end;
    xcd_chunk_aligned<4>(n, beg, end);
    constexpr bool kMinv = !SCALED || RR;
    double rr = 0.0, rz = 0.0;
    auto body = [&](float& ri, float api, float mi) {
        ri = ri - alpha * api;
        if constexpr (SCALED)
        {
            double const q = double(ri) * double(ri);
            rz += q;
            if constexpr (RR) rr += q / double(mi);
        }
        else
        {
            float const zi = mi * ri;
            rr += double(ri) * double(ri);
            rz += double(ri) * double(zi);
        }
    };
    // four entries per thread in 16-byte accesses, then the chunk's tail
    int64_t const full = beg + (end - beg) / 4 * 4;
    for (int64_t i = beg + 4 * threadIdx.x; i < full; i += 4 * kBlock)
    {
        float4v rv = *reinterpret_cast<const float4v*>(r32 + i);
        float4v const av = load_stream<kVectorNT>(reinterpret_cast<const float4v*>(Ap32 + i));
        float4v mv{1.0f, 1.0f, 1.0f, 1.0f};
        if constexpr (kMinv) mv = *reinterpret_cast<const float4v*>(minv + i);
        float r0 = rv.x, r1 = rv.y, r2 = rv.z, r3 = rv.w;
        body(r0, av.x, mv.x);
        body(r1, av.y, mv.y);
        body(r2, av.z, mv.z);
        body(r3, av.w, mv.w);
        *reinterpret_cast<float4v*>(r32 + i) = float4v{r0, r1, r2, r3};
    }
    for (int64_t i = full + threadIdx.x; i < end; i += kBlock)
    {
        float ri = r32[i];
        body(ri, Ap32[i], kMinv ? minv[i] : 1.0f);
        r32[i] = ri;
    }
    if constexpr (kMinv) rr = block_sum<kBlock>(rr);
    rz = block_sum<kBlock>(rz);
    if (threadIdx.x == 0)
    {
        if constexpr (kMinv) part_rr[blockIdx.x] = rr;
        part_rz[blockIdx.x] = rz;
    }
}

// single block: decide whether this iteration ends with a reliable update
__global__ __launch_bounds__(kBlock) void mixed_decide_kernel(const double* __restrict__ part_rr,
                                                              const double* __restrict__ part_rz, int npart,
                                                              krylov_state* __restrict__ st, double delta)
{
    if (st->done) return;
    double const rr = partials_total<kBlock>(part_rr, npart);
    double const rz = partials_total<kBlock>(part_rz, npart);
    if (threadIdx.x != 0) return;
    bool const refine = rr < st->threshold || sqrt(rr) < delta * st->aux0 || !isfinite(rr) || !isfinite(rz);
    st->refine = refine ? 1 : 0;
    st->rr = rr;
    st->aux1 = rz;
}

// x64 += x32 + alpha p32, x32 = 0 when a reliable update is due (the iteration's x step, which
// the p update then skips); with force, x64 += x32, x32 = 0 unconditionally (the recurrence's
// remainder after the loop).  With minv64 the recurrence is that of the Jacobi-scaled system,
// so its step is scaled back: x64 += sqrt(minv64) (x32 + alpha p32)
__global__ __launch_bounds__(kBlock) void mixed_fold_kernel(int64_t n, double* __restrict__ x64,
                                                            float* __restrict__ x32,
                                                            const float* __restrict__ p32,
                                                            const double* __restrict__ part_pAp,
                                                            const krylov_state* __restrict__ st, int force,
                                                            const double* __restrict__ minv64)
{
    if (!force && (st->done || !st->refine)) return;
    float const alpha = force ? 0.0f : float(st->rz / partials_total<kBlock>(part_pAp, -1));
    int64_t beg, end;
    xcd_chunk(n, beg, end);
    for (int64_t i = beg + threadIdx.x; i < end; i += kBlock)
    {
        float const xi = force ? x32[i] : x32[i] + alpha * p32[i];
        x64[i] += minv64 ? sqrt(minv64[i]) * double(xi) : double(xi);
        x32[i] = 0.0f;
    }
}

// r32 = float(r64) after the FP64 residual; partials of r64.r64 and r32.z32.  SCALED: the
// recurrence holds the scaled residual r32 = float(D^-1/2 r64), and z~ = r~
template <bool SCALED>
__global__ __launch_bounds__(kBlock) void mixed_refresh_kernel(int64_t n, const double* __restrict__ r64,
                                                               float* __restrict__ r32,
                                                               const float* __restrict__ minv,
                                                               const double* __restrict__ minv64,
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
        rr += r * r;
        if constexpr (SCALED)
        {
            float const ri = float(sqrt(minv64[i]) * r);
            r32[i] = ri;
            rz += double(ri) * double(ri);
        }
        else
        {
            float const ri = float(r);
            r32[i] = ri;
            rz += double(ri) * double(minv[i] * ri);
        }
    }
    rr = block_sum<kBlock>(rr);
    rz = block_sum<kBlock>(rz);
    if (threadIdx.x == 0)
    {
        part_rr[blockIdx.x] = rr;
        part_rz[blockIdx.x] = rz;
    }
}

// x32 += alpha p32 (unless a reliable update folded it already), convergence on the true
// residual, beta, p32 = z + beta p32; block 0 writes the next state (decided = 0: no decision
// kernel ran this iteration, so r.r and r.z come from the partials of the x/r update and no
// reliable update is due).  SCALED: z = r (no minv read), and without a decision only r.z is
// reduced (r.r is read only by the decision)
template <bool SCALED>
__global__ __launch_bounds__(kBlock) void mixed_update_p_kernel(int64_t n, float* __restrict__ p32,
                                                                float* __restrict__ x32,
                                                                const float* __restrict__ r32,
                                                                const float* __restrict__ minv,
                                                                const double* __restrict__ part_pAp,
                                                                const double* __restrict__ part_rr_true,
                                                                const double* __restrict__ part_rz_true,
                                                                const double* __restrict__ part_rr,
                                                                const double* __restrict__ part_rz, int npart,
                                                                int npart_true, int decided,
                                                                const krylov_state* __restrict__ st,
                                                                krylov_state* __restrict__ st_next, int max_it)
{
    krylov_state s = *st;
    if (s.done)
    {
        if (blockIdx.x == 0 && threadIdx.x == 0) *st_next = s;
        return;
    }
    float const alpha = float(s.rz / partials_total<kBlock>(part_pAp, -1));
    double rr = s.rr, rz = s.aux1;
    if (!decided)
    {
        if constexpr (!SCALED) rr = partials_total<kBlock>(part_rr, npart);
        rz = partials_total<kBlock>(part_rz, npart);
        s.refine = 0;
    }
    bool const step_x = !s.refine;
    bool conv = false;
    if (s.refine)
    {
        rr = reduce_partials<kBlock>(part_rr_true, npart_true);
        rz = reduce_partials<kBlock>(part_rz_true, npart_true);
        conv = rr < s.threshold;
        s.aux0 = sqrt(rr);
        s.restarts += 1; // counts reliable updates
    }
    int32_t const it_new = conv ? s.it : s.it + 1;
    int32_t done = conv ? KRYLOV_CONVERGED : (it_new >= max_it ? KRYLOV_LIMIT : KRYLOV_RUNNING);
    if (!isfinite(rr) || !isfinite(rz)) done = KRYLOV_NONFINITE;
    float const beta = float(rz / s.rz);
    bool const more = done == KRYLOV_RUNNING;
    if (step_x || more)
    {
        int64_t beg, end;
        xcd_chunk_aligned<4>(n, beg, end);
        int64_t const full = beg + (end - beg) / 4 * 4;
        for (int64_t i = beg + 4 * threadIdx.x; i < full; i += 4 * kBlock)
        {
            float4v const pv = *reinterpret_cast<const float4v*>(p32 + i);
            if (step_x)
            {
                float4v const xv = load_stream<kVectorNT>(reinterpret_cast<const float4v*>(x32 + i));
                store_stream<kVectorNT>(reinterpret_cast<float4v*>(x32 + i),
                                        float4v{xv.x + alpha * pv.x, xv.y + alpha * pv.y, xv.z + alpha * pv.z,
                                                xv.w + alpha * pv.w});
            }
            if (more)
            {
                float4v const rv = *reinterpret_cast<const float4v*>(r32 + i);
                if constexpr (SCALED)
                {
                    *reinterpret_cast<float4v*>(p32 + i) =
                        float4v{rv.x + beta * pv.x, rv.y + beta * pv.y, rv.z + beta * pv.z, rv.w + beta * pv.w};
                }
                else
                {
                    float4v const mv = *reinterpret_cast<const float4v*>(minv + i);
                    *reinterpret_cast<float4v*>(p32 + i) =
                        float4v{mv.x * rv.x + beta * pv.x, mv.y * rv.y + beta * pv.y, mv.z * rv.z + beta * pv.z,
                                mv.w * rv.w + beta * pv.w};
                }
            }
        }
        for (int64_t i = full + threadIdx.x; i < end; i += kBlock)
        {
            float const pi = p32[i];
            if (step_x) x32[i] = x32[i] + alpha * pi;
            if (more) p32[i] = (SCALED ? r32[i] : minv[i] * r32[i]) + beta * pi;
        }
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
