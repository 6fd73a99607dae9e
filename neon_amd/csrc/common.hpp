// Shared device helpers and host-side error handling for the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace neon_amd
{
// A fatal or recoverable status travelling through host code as an exception; the C ABI
// turns it back into the int status of neon_hip.h.
struct status_error : std::runtime_error
{
    int code;
    status_error(int c, std::string const& what) : std::runtime_error(what), code(c) {}
};

#define NEON_CHECK_HIP(expr)                                                                           \
    do                                                                                                 \
    {                                                                                                  \
        hipError_t const e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                          \
            throw ::neon_amd::status_error(e_ == hipErrorOutOfMemory ? NEON_OOM : NEON_HIP_ERROR,      \
                                           std::string(#expr) + ": " + hipGetErrorString(e_));        \
    } while (0)

// Device status word bits (atomic OR from kernels, read by the host after a phase).
enum : uint32_t
{
    STATUS_NEGATIVE_JACOBIAN = 1u,
    STATUS_NAN = 2u,
    STATUS_RETURN_MAP = 4u,
};

constexpr int kBlock = 256;        // threads per block for vector/SpMV kernels
#ifndef NEON_MAX_PARTIALS
#define NEON_MAX_PARTIALS 16384
#endif
constexpr int kMaxPartials = NEON_MAX_PARTIALS; // partial-sum slots for grid reductions

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over a block of BS threads; every thread receives the total.  Fixed order.
template <int BS>
__device__ inline double block_sum(double v)
{
    __shared__ double sh[BS / 64];
    v = wave_sum(v);
    int const w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < BS / 64; ++k) t += sh[k];
    __syncthreads();
    return t;
}

// Every block reduces the partials written by the previous kernel (deterministic,
// identical in every block, so no extra launch or cross-block handshake is needed).
template <int BS>
__device__ inline double reduce_partials(const double* __restrict__ part, int n)
{
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += BS) s += part[k];
    return block_sum<BS>(s);
}

// The total of a partials array: n >= 0 reduces its n partials (as reduce_partials); n < 0
// reads the total that totals_kernel stored after them, at part[kMaxPartials].
template <int BS>
__device__ inline double partials_total(const double* __restrict__ part, int n)
{
    if (n < 0) return part[kMaxPartials];
    return reduce_partials<BS>(part, n);
}

// Contiguous [begin, end) chunk of `total` items for a block, with the blocks of each
// XCD (block id mod 8 on MI355X) given adjacent chunks so that neighbouring rows share
// an L2.
__device__ inline void xcd_chunk(int64_t total, int64_t& begin, int64_t& end)
{
    int const nb = gridDim.x;
    int const b = blockIdx.x;
    int logical = b;
    if ((nb & 7) == 0) logical = (b & 7) * (nb >> 3) + (b >> 3);
    int64_t const per = (total + nb - 1) / nb;
    begin = int64_t(logical) * per;
    end = begin + per < total ? begin + per : total;
}
// The same split with chunk boundaries at multiples of A items (the last chunk ends at
// total), so that a chunk of an A-aligned array starts on an A-item boundary.
template <int A>
__device__ inline void xcd_chunk_aligned(int64_t total, int64_t& begin, int64_t& end)
{
    int const nb = gridDim.x;
    int const b = blockIdx.x;
    int logical = b;
    if ((nb & 7) == 0) logical = (b & 7) * (nb >> 3) + (b >> 3);
    int64_t const units = (total + A - 1) / A;
    int64_t const per = (units + nb - 1) / nb;
    int64_t const b0 = int64_t(logical) * per * A;
    begin = b0 < total ? b0 : total;
    end = begin + per * A < total ? begin + per * A : total;
}

typedef float float4v __attribute__((ext_vector_type(4)));
typedef double double2v __attribute__((ext_vector_type(2)));
} // namespace neon_amd
