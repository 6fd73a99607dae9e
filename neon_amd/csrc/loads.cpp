// Load vectors of unit nonfollower loads on the initial configuration (host only).
//
// neon integrates its surface and volume loads over the undeformed geometry
// (neumann.hpp:104-118, 159-173; pressure.cpp:8-31), so a load is its history value times a
// fixed vector; these functions build that vector once for the Newton driver.
//   quadrilateral4: node a at (xi_a, eta_a) = (-1,-1) (1,-1) (1,1) (-1,1),
//     N_a = (1 + xi_a xi)(1 + eta_a eta) / 4 (quadrilateral.cpp:17-46); 2x2 Gauss points of
//     weight 1, or the centre with weight 4 when reduced (quadrilateral_quadrature.cpp:10-24);
//     surface measure j = |x_xi x x_eta| (jacobian_determinant.hpp:19-22).
//   hexahedron8: the local node order of hexahedron.cpp:21-28, 2x2x2 Gauss points of weight
//     1, or the centre with weight 8 when reduced; measure det J.
#include "neon_hip.h"

#include <cmath>
#include <cstdint>

namespace
{
constexpr double kGauss = 0.57735026918962576451; // 1 / sqrt(3)
constexpr double kQuadXi[4] = {-1.0, 1.0, 1.0, -1.0};
constexpr double kQuadEta[4] = {-1.0, -1.0, 1.0, 1.0};
constexpr double kHexXi[8] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0};
constexpr double kHexEta[8] = {-1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0};
constexpr double kHexZeta[8] = {-1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0};

bool nodes_in_range(int64_t count, const int32_t* idx, int32_t n_nodes)
{
    for (int64_t i = 0; i < count; ++i)
        if (idx[i] < 0 || idx[i] >= n_nodes) return false;
    return true;
}
} // namespace

extern "C" {

int neon_surface_load_vector(int32_t n_nodes, const double* X, int32_t n_faces, const int32_t* faces,
                             int32_t component, int32_t reduced, double* out)
{
    if (n_nodes <= 0 || !X || n_faces < 0 || (n_faces > 0 && !faces) || !out || component < -1 || component > 2)
        return NEON_BAD_ARGUMENT;
    if (!nodes_in_range(4ll * n_faces, faces, n_nodes)) return NEON_BAD_ARGUMENT;
    int const points = reduced ? 1 : 4;
    for (int32_t f = 0; f < n_faces; ++f)
    {
        const int32_t* nd = faces + 4ll * f;
        for (int l = 0; l < points; ++l)
        {
            double const xi = reduced ? 0.0 : kGauss * kQuadXi[l];
            double const eta = reduced ? 0.0 : kGauss * kQuadEta[l];
            double const w = reduced ? 4.0 : 1.0;
            double N[4], t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
            for (int a = 0; a < 4; ++a)
            {
                N[a] = 0.25 * (1.0 + kQuadXi[a] * xi) * (1.0 + kQuadEta[a] * eta);
                double const dxi = 0.25 * kQuadXi[a] * (1.0 + kQuadEta[a] * eta);
                double const deta = 0.25 * kQuadEta[a] * (1.0 + kQuadXi[a] * xi);
                for (int i = 0; i < 3; ++i)
                {
                    t1[i] += X[3ll * nd[a] + i] * dxi;
                    t2[i] += X[3ll * nd[a] + i] * deta;
                }
            }
            double const c[3] = {t1[1] * t2[2] - t1[2] * t2[1], t1[2] * t2[0] - t1[0] * t2[2],
                                 t1[0] * t2[1] - t1[1] * t2[0]};
            double const j = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            for (int a = 0; a < 4; ++a)
            {
                if (component >= 0)
                    out[3ll * nd[a] + component] += w * N[a] * j;
                else
                    // pressure: -N_a n j with n = c / j
                    for (int i = 0; i < 3; ++i) out[3ll * nd[a] + i] -= w * N[a] * c[i];
            }
        }
    }
    return NEON_OK;
}

int neon_volume_load_vector(int32_t n_nodes, const double* X, int32_t n_elem, const int32_t* conn,
                            int32_t component, int32_t reduced, double* out)
{
    if (n_nodes <= 0 || !X || n_elem < 0 || (n_elem > 0 && !conn) || !out || component < 0 || component > 2)
        return NEON_BAD_ARGUMENT;
    if (!nodes_in_range(8ll * n_elem, conn, n_nodes)) return NEON_BAD_ARGUMENT;
    int const points = reduced ? 1 : 8;
    for (int32_t e = 0; e < n_elem; ++e)
    {
        const int32_t* nd = conn + 8ll * e;
        for (int l = 0; l < points; ++l)
        {
            double const xi = reduced ? 0.0 : kGauss * kHexXi[l];
            double const eta = reduced ? 0.0 : kGauss * kHexEta[l];
            double const zeta = reduced ? 0.0 : kGauss * kHexZeta[l];
            double const w = reduced ? 8.0 : 1.0;
            double N[8], J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            for (int a = 0; a < 8; ++a)
            {
                double const fx = 1.0 + kHexXi[a] * xi, fy = 1.0 + kHexEta[a] * eta, fz = 1.0 + kHexZeta[a] * zeta;
                N[a] = 0.125 * fx * fy * fz;
                double const d[3] = {0.125 * kHexXi[a] * fy * fz, 0.125 * kHexEta[a] * fx * fz,
                                     0.125 * kHexZeta[a] * fx * fy};
                for (int i = 0; i < 3; ++i)
                    for (int k = 0; k < 3; ++k) J[i][k] += X[3ll * nd[a] + i] * d[k];
            }
            double const det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
                               - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
                               + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            for (int a = 0; a < 8; ++a) out[3ll * nd[a] + component] += w * N[a] * det;
        }
    }
    return NEON_OK;
}

} // extern "C"
