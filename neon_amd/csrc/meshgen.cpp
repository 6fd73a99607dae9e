// Synthetic structured hex8 meshes and counter-based random streams (host only).
//
// Node numbering n = i + (nx+1)(j + (ny+1)k) and hex8 connectivity in the local node order of
// src/interpolations/hexahedron.cpp:21-28: (i,j,k) (i+1,j,k) (i+1,j+1,k) (i,j+1,k) and the
// same on layer k+1.
//
// Random numbers: SplitMix64 (G. Steele, D. Lea, C. Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014), used counter-style: value i of stream `seed` is the
// (i+1)-th output of the sequential generator seeded with `seed`, so any index can be
// produced independently and in parallel.
#include "neon_hip.h"

#include <cmath>
#include <cstdint>
#include <vector>

namespace
{
inline uint64_t splitmix64_at(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double uniform01(uint64_t seed, uint64_t i) { return double(splitmix64_at(seed, i) >> 11) * 0x1.0p-53; }

struct box
{
    int32_t nx, ny, nz;
    double Lx, Ly, Lz, r;
    bool keep(int32_t i, int32_t j, int32_t k) const
    {
        if (r <= 0.0) return true;
        double const hx = Lx / nx, hy = Ly / ny;
        double const cx = (i + 0.5) * hx - 0.5 * Lx, cy = (j + 0.5) * hy - 0.5 * Ly;
        (void)k;
        return cx * cx + cy * cy > r * r;
    }
};

bool valid(int32_t nx, int32_t ny, int32_t nz, double Lx, double Ly, double Lz)
{
    return nx > 0 && ny > 0 && nz > 0 && Lx > 0 && Ly > 0 && Lz > 0
           && int64_t(nx + 1) * (ny + 1) * (nz + 1) < int64_t(INT32_MAX) / 3;
}
} // namespace

extern "C" {

int neon_random_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out)
{
    if (n < 0 || (n > 0 && !out)) return NEON_BAD_ARGUMENT;
    for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * uniform01(seed, uint64_t(i));
    return NEON_OK;
}

int neon_random_normal(uint64_t seed, int64_t n, double* out)
{
    if (n < 0 || (n > 0 && !out)) return NEON_BAD_ARGUMENT;
    double const two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < n; ++i)
    {
        double const u1 = uniform01(seed, 2 * uint64_t(i)), u2 = uniform01(seed, 2 * uint64_t(i) + 1);
        out[i] = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(two_pi * u2);
    }
    return NEON_OK;
}

int neon_mesh_box_size(int32_t nx, int32_t ny, int32_t nz, double Lx, double Ly, double Lz, double hole_radius,
                       int32_t* n_nodes, int32_t* n_elem)
{
    if (!valid(nx, ny, nz, Lx, Ly, Lz) || !n_nodes || !n_elem) return NEON_BAD_ARGUMENT;
    box const B{nx, ny, nz, Lx, Ly, Lz, hole_radius};
    int64_t const NX = nx + 1, NY = ny + 1, NZ = nz + 1;
    std::vector<uint8_t> used(size_t(NX * NY * NZ), 0);
    int64_t ne = 0;
    for (int32_t k = 0; k < nz; ++k)
        for (int32_t j = 0; j < ny; ++j)
            for (int32_t i = 0; i < nx; ++i)
                if (B.keep(i, j, k))
                {
                    ++ne;
                    for (int c = 0; c < 8; ++c)
                    {
                        int const di = (c == 1 || c == 2 || c == 5 || c == 6), dj = (c == 2 || c == 3 || c == 6 || c == 7),
                                  dk = c >= 4;
                        used[(i + di) + NX * ((j + dj) + NY * (k + dk))] = 1;
                    }
                }
    int64_t nn = 0;
    for (auto u : used) nn += u;
    *n_nodes = int32_t(nn);
    *n_elem = int32_t(ne);
    return NEON_OK;
}

int neon_mesh_box(int32_t nx, int32_t ny, int32_t nz, double Lx, double Ly, double Lz, double jitter, uint64_t seed,
                  double hole_radius, double* X, int32_t* conn, int32_t* node_ijk)
{
    if (!valid(nx, ny, nz, Lx, Ly, Lz) || !X || !conn) return NEON_BAD_ARGUMENT;
    box const B{nx, ny, nz, Lx, Ly, Lz, hole_radius};
    int64_t const NX = nx + 1, NY = ny + 1, NZ = nz + 1;
    std::vector<uint8_t> used(size_t(NX * NY * NZ), 0);
    for (int32_t k = 0; k < nz; ++k)
        for (int32_t j = 0; j < ny; ++j)
            for (int32_t i = 0; i < nx; ++i)
                if (B.keep(i, j, k))
                    for (int c = 0; c < 8; ++c)
                    {
                        int const di = (c == 1 || c == 2 || c == 5 || c == 6), dj = (c == 2 || c == 3 || c == 6 || c == 7),
                                  dk = c >= 4;
                        used[(i + di) + NX * ((j + dj) + NY * (k + dk))] = 1;
                    }
    std::vector<int32_t> id(used.size(), -1);
    int32_t nn = 0;
    for (size_t g = 0; g < used.size(); ++g)
        if (used[g]) id[g] = nn++;
    double const h[3] = {Lx / nx, Ly / ny, Lz / nz};
    int32_t const nmax[3] = {nx, ny, nz};
    for (int64_t k = 0; k < NZ; ++k)
        for (int64_t j = 0; j < NY; ++j)
            for (int64_t i = 0; i < NX; ++i)
            {
                int64_t const g = i + NX * (j + NY * k);
                if (id[g] < 0) continue;
                int64_t const idx[3] = {i, j, k};
                double* x = X + 3ll * id[g];
                for (int c = 0; c < 3; ++c)
                {
                    double v = idx[c] * h[c];
                    bool const on_face = idx[c] == 0 || idx[c] == nmax[c];
                    if (!on_face && jitter != 0.0) v += (2.0 * uniform01(seed, uint64_t(3 * g + c)) - 1.0) * jitter * h[c];
                    x[c] = v;
                }
                if (node_ijk)
                {
                    node_ijk[3ll * id[g]] = int32_t(i);
                    node_ijk[3ll * id[g] + 1] = int32_t(j);
                    node_ijk[3ll * id[g] + 2] = int32_t(k);
                }
            }
    int64_t e = 0;
    for (int32_t k = 0; k < nz; ++k)
        for (int32_t j = 0; j < ny; ++j)
            for (int32_t i = 0; i < nx; ++i)
                if (B.keep(i, j, k))
                {
                    for (int c = 0; c < 8; ++c)
                    {
                        int const di = (c == 1 || c == 2 || c == 5 || c == 6), dj = (c == 2 || c == 3 || c == 6 || c == 7),
                                  dk = c >= 4;
                        conn[8 * e + c] = id[(i + di) + NX * ((j + dj) + NY * (k + dk))];
                    }
                    ++e;
                }
    return NEON_OK;
}

} // extern "C"
