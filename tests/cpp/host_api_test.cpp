// C++ host classes of include/neon_hip.hpp on the GPU, against results that the mathematics
// fixes:
//   * the tridiagonal system of the reference's linear-solver test (test/linear_solvers.cpp:15-49),
//     x = (1.375, 1.75, 1.625), through make_linear_solver's conjugate gradient and BiCGSTAB;
//   * factory errors for unknown options;
//   * a uniaxial-stress patch test: a jittered neo-Hooke box on rollers stretched by s along z
//     deforms homogeneously, u = (F - I) X with F = diag(l, l, 1 + s), where the lateral stretch
//     l solves lambda ln J + mu (l^2 - 1) = 0, J = l^2 (1 + s) (sigma_xx = 0 for
//     compressible_neohooke.cpp:37-40); hex8 interpolates linear fields exactly, so the FE
//     solution equals it to solver accuracy.
// Exit status 0: all checks passed; 77: no HIP device (nothing run); 1: a check failed.
#include "neon_hip.hpp"

#include <cmath>
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

namespace
{
int failures = 0;

void expect(bool ok, char const* what)
{
    std::printf("%s: %s\n", ok ? "ok  " : "FAIL", what);
    if (!ok) ++failures;
}

void tridiagonal(neon_hip::context& ctx)
{
    std::vector<std::int32_t> rp{0, 2, 5, 7}, col{0, 1, 0, 1, 2, 1, 2};
    std::vector<double> val{2.0, -1.0, -1.0, 2.0, -1.0, -1.0, 2.0}, b{1.0, 0.5, 1.5}, x;
    neon_hip::csr_view A{3, 7, rp.data(), col.data(), val.data()};
    neon_hip::solver_options o;
    o.type = "iterative";
    o.device = "gpu";
    o.backend = "hip";
    o.tolerance = 1e-12;
    for (bool symmetric : {true, false})
    {
        auto s = neon_hip::make_linear_solver(o, symmetric);
        auto* it = dynamic_cast<neon_hip::iterative_linear_solver*>(s.get());
        it->ctx = &ctx;
        s->solve(A, x, b);
        bool const close = std::abs(x[0] - 1.375) < 1e-10 && std::abs(x[1] - 1.75) < 1e-10
                           && std::abs(x[2] - 1.625) < 1e-10;
        expect(close, symmetric ? "tridiagonal golden system, conjugate gradient" : "tridiagonal golden system, BiCGSTAB");
        expect(symmetric == (dynamic_cast<neon_hip::conjugate_gradient_hip*>(s.get()) != nullptr),
               "factory picks CG for symmetric and BiCGSTAB otherwise");
    }
    // a solve that cannot converge in 1 iteration is recoverable
    neon_hip::conjugate_gradient_hip one(1e-14, 1);
    one.ctx = &ctx;
    one.criterion = NEON_STOP_RELATIVE;
    bool recoverable = false;
    try
    {
        one.solve(A, x, b);
    }
    catch (neon_hip::computational_error const& e)
    {
        recoverable = e.code == NEON_NONCONVERGED;
    }
    expect(recoverable, "iteration limit throws computational_error");
}

void factory_errors()
{
    neon_hip::solver_options o;
    bool threw = false;
    try
    {
        neon_hip::make_linear_solver(o, true);
    }
    catch (std::domain_error const&)
    {
        threw = true;
    }
    expect(threw, "missing solver type throws domain_error");
    o.type = "iterative";
    o.device = "gpu";
    o.backend = "cuda";
    threw = false;
    try
    {
        neon_hip::make_linear_solver(o, true);
    }
    catch (std::domain_error const&)
    {
        threw = true;
    }
    expect(threw, "unknown backend throws domain_error");
}

double lateral_stretch(double bulk, double shear, double stretch)
{
    double const lambda = bulk - 2.0 * shear / 3.0;
    double l = 1.0;
    for (int k = 0; k < 50; ++k)
    {
        double const J = l * l * stretch;
        double const r = lambda * std::log(J) + shear * (l * l - 1.0);
        double const dr = lambda * 2.0 / l + 2.0 * shear * l;
        l -= r / dr;
    }
    return l;
}

void patch_test(neon_hip::context& ctx)
{
    int const n = 6;
    double const s = 0.05, bulk = 2.0, shear = 0.4;
    std::int32_t nn = 0, ne = 0;
    neon_mesh_box_size(n, n, n, 1.0, 1.0, 1.0, 0.0, &nn, &ne);
    std::vector<double> X(3 * std::size_t(nn));
    std::vector<std::int32_t> conn(8 * std::size_t(ne)), ijk(3 * std::size_t(nn));
    neon_mesh_box(n, n, n, 1.0, 1.0, 1.0, 0.1, 0x6E656F6E01ull, 0.0, X.data(), conn.data(), ijk.data());
    std::vector<double> params(4 * std::size_t(ne));
    for (std::int32_t e = 0; e < ne; ++e)
    {
        params[4 * e] = bulk;
        params[4 * e + 1] = shear;
    }
    neon_hip::device_mesh mesh(ctx, X, conn, params, NEON_MODEL_NEOHOOKE);
    std::vector<neon_hip::dirichlet_boundary> bcs(4);
    for (std::int32_t a = 0; a < nn; ++a)
    {
        if (ijk[3 * a] == 0) bcs[0].dofs.push_back(3 * a);
        if (ijk[3 * a + 1] == 0) bcs[1].dofs.push_back(3 * a + 1);
        if (ijk[3 * a + 2] == 0) bcs[2].dofs.push_back(3 * a + 2);
        if (ijk[3 * a + 2] == n) bcs[3].dofs.push_back(3 * a + 2);
    }
    for (int k = 0; k < 4; ++k)
    {
        bcs[k].times = {0.0, 1.0};
        bcs[k].values = {0.0, k == 3 ? s : 0.0};
    }
    neon_hip::simulation_options o;
    o.initial = o.minimum = o.maximum = 0.5;
    o.residual_tolerance = 1e-10;
    o.displacement_tolerance = 1e-10;
    o.always_relative = true;
    o.linear.type = "iterative";
    o.linear.device = "gpu";
    o.linear.backend = "hip";
    o.linear.tolerance = 1e-13;
    o.linear.maximum_iterations = 5000;
    neon_hip::static_matrix sm(mesh, o, bcs);
    std::vector<double> const u = sm.solve();
    double const l = lateral_stretch(bulk, shear, 1.0 + s);
    double err = 0.0, umax = 0.0;
    for (std::int32_t a = 0; a < nn; ++a)
    {
        double const ue[3] = {(l - 1.0) * X[3 * a], (l - 1.0) * X[3 * a + 1], s * X[3 * a + 2]};
        for (int i = 0; i < 3; ++i)
        {
            err = std::max(err, std::abs(u[3 * a + i] - ue[i]));
            umax = std::max(umax, std::abs(ue[i]));
        }
    }
    std::printf("      patch test: lateral stretch %.12f, max |u - (F - I) X| / max|u| = %.3e, %zu Newton records\n", l,
                err / umax, sm.records().size());
    expect(err <= 1e-9 * umax, "uniaxial neo-Hooke patch test matches the closed form");
    expect(sm.records().size() >= 2, "two load steps recorded");

    // LATIN with the zero search direction on a fresh mesh: the displacement-driven stretch
    // reaches the same state; without a direction latin_matrix refuses, as neon's constructor
    bool refused = false;
    try
    {
        neon_hip::latin_matrix bad(mesh, o, bcs);
    }
    catch (std::domain_error const& e)
    {
        refused = std::string(e.what()).find("latin_search_direction") != std::string::npos;
    }
    expect(refused, "latin_matrix needs latin_search_direction");
    neon_hip::device_mesh mesh2(ctx, X, conn, params, NEON_MODEL_NEOHOOKE);
    neon_hip::simulation_options ol = o;
    ol.latin_search_direction = 0.0;
    neon_hip::latin_matrix lm(mesh2, ol, bcs);
    std::vector<double> const ul = lm.solve();
    double dl = 0.0;
    for (std::size_t i = 0; i < u.size(); ++i) dl = std::max(dl, std::abs(ul[i] - u[i]));
    std::printf("      LATIN (s = 0): max |u_latin - u_newton| / max|u| = %.3e, %zu records\n", dl / umax,
                lm.records().size());
    expect(dl <= 1e-9 * umax, "LATIN with s = 0 reaches the Newton solution of the patch test");
}
// unit-load vectors of the box's top face (area 1) and volume (1): host-only calls
void load_vectors()
{
    int const n = 4;
    std::int32_t nn = 0, ne = 0;
    neon_mesh_box_size(n, n, n, 1.0, 1.0, 1.0, 0.0, &nn, &ne);
    std::vector<double> X(3 * std::size_t(nn));
    std::vector<std::int32_t> conn(8 * std::size_t(ne)), ijk(3 * std::size_t(nn));
    neon_mesh_box(n, n, n, 1.0, 1.0, 1.0, 0.1, 0x6E656F6E01ull, 0.0, X.data(), conn.data(), ijk.data());
    std::vector<std::int32_t> top;
    for (std::int32_t e = 0; e < ne; ++e)
        if (ijk[3 * std::size_t(conn[8 * e + 4]) + 2] == n) top.insert(top.end(), &conn[8 * e + 4], &conn[8 * e + 8]);
    std::vector<double> const t = neon_hip::surface_load_vector(X, top, 2);
    std::vector<double> const p = neon_hip::surface_load_vector(X, top, -1);
    std::vector<double> const b = neon_hip::volume_load_vector(X, conn, 1);
    double tz = 0.0, pz = 0.0, by = 0.0, other = 0.0;
    for (std::int32_t a = 0; a < nn; ++a)
    {
        tz += t[3 * a + 2];
        pz += p[3 * a + 2];
        by += b[3 * a + 1];
        other += std::abs(t[3 * a]) + std::abs(t[3 * a + 1]) + std::abs(b[3 * a]) + std::abs(b[3 * a + 2]);
    }
    std::printf("      load vectors: traction %.15f, pressure %.15f, body force %.15f\n", tz, pz, by);
    expect(std::abs(tz - 1.0) < 1e-13 && std::abs(pz + 1.0) < 1e-13 && std::abs(by - 1.0) < 1e-13 && other == 0.0,
           "unit-load vectors integrate the face area and the volume");
    bool threw = false;
    try
    {
        neon_hip::volume_load_vector(X, conn, 3);
    }
    catch (std::domain_error const&)
    {
        threw = true;
    }
    expect(threw, "a body force component outside 0..2 is rejected");
}

// uniaxial neo-Hooke under a top traction t (first Piola stress P_zz = t, no lateral stress):
// lambda ln J + mu (l^2 - 1) = 0 and (lambda ln J + mu (lz^2 - 1)) / lz = t, J = l^2 lz
void traction_stretches(double bulk, double shear, double t, double& lz, double& l)
{
    double const lambda = bulk - 2.0 * shear / 3.0;
    lz = 1.0;
    l = 1.0;
    for (int it = 0; it < 60; ++it)
    {
        double const lnJ = std::log(l * l * lz);
        double const r1 = lambda * lnJ + shear * (l * l - 1.0);
        double const r2 = (lambda * lnJ + shear * (lz * lz - 1.0)) / lz - t;
        double const a11 = 2.0 * lambda / l + 2.0 * shear * l, a12 = lambda / lz;
        double const a21 = 2.0 * lambda / (l * lz);
        double const a22 = (lambda / lz + 2.0 * shear * lz) / lz - (lambda * lnJ + shear * (lz * lz - 1.0)) / (lz * lz);
        double const det = a11 * a22 - a12 * a21;
        l -= (r1 * a22 - a12 * r2) / det;
        lz -= (a11 * r2 - a21 * r1) / det;
    }
}

void traction_patch_test(neon_hip::context& ctx)
{
    int const n = 4;
    double const t = 0.1, bulk = 2.0, shear = 0.4;
    std::int32_t nn = 0, ne = 0;
    neon_mesh_box_size(n, n, n, 1.0, 1.0, 1.0, 0.0, &nn, &ne);
    std::vector<double> X(3 * std::size_t(nn));
    std::vector<std::int32_t> conn(8 * std::size_t(ne)), ijk(3 * std::size_t(nn));
    neon_mesh_box(n, n, n, 1.0, 1.0, 1.0, 0.1, 0x6E656F6E01ull, 0.0, X.data(), conn.data(), ijk.data());
    std::vector<double> params(4 * std::size_t(ne));
    for (std::int32_t e = 0; e < ne; ++e)
    {
        params[4 * e] = bulk;
        params[4 * e + 1] = shear;
    }
    neon_hip::device_mesh mesh(ctx, X, conn, params, NEON_MODEL_NEOHOOKE);
    std::vector<neon_hip::dirichlet_boundary> bcs(3);
    for (std::int32_t a = 0; a < nn; ++a)
        for (int i = 0; i < 3; ++i)
            if (ijk[3 * a + i] == 0) bcs[i].dofs.push_back(3 * a + i);
    for (auto& b : bcs)
    {
        b.times = {0.0, 1.0};
        b.values = {0.0, 0.0};
    }
    std::vector<std::int32_t> top;
    for (std::int32_t e = 0; e < ne; ++e)
        if (ijk[3 * std::size_t(conn[8 * e + 4]) + 2] == n) top.insert(top.end(), &conn[8 * e + 4], &conn[8 * e + 8]);
    std::vector<double> const unit = neon_hip::surface_load_vector(X, top, 2);
    neon_hip::nonfollower_load load;
    for (std::int32_t i = 0; i < std::int32_t(unit.size()); ++i)
        if (unit[i] != 0.0)
        {
            load.dofs.push_back(i);
            load.values.push_back(unit[i]);
        }
    load.times = {0.0, 1.0};
    load.loads = {0.0, t};
    neon_hip::simulation_options o;
    o.initial = o.minimum = o.maximum = 0.5;
    o.residual_tolerance = 1e-10;
    o.displacement_tolerance = 1e-10;
    o.always_relative = true;
    o.linear.type = "iterative";
    o.linear.device = "gpu";
    o.linear.backend = "hip";
    o.linear.tolerance = 1e-13;
    o.linear.maximum_iterations = 5000;
    neon_hip::static_matrix sm(mesh, o, bcs, {load});
    std::vector<double> hook_times;
    sm.on_step = [&](std::int32_t, double time) { hook_times.push_back(time); };
    std::vector<double> const u = sm.solve();
    expect(hook_times.size() == 3 && hook_times.front() == 0.0 && hook_times.back() == 1.0,
           "the step hook sees the initial state and both converged steps");
    double lz = 1.0, l = 1.0;
    traction_stretches(bulk, shear, t, lz, l);
    double err = 0.0, umax = 0.0;
    for (std::int32_t a = 0; a < nn; ++a)
    {
        double const ue[3] = {(l - 1.0) * X[3 * a], (l - 1.0) * X[3 * a + 1], (lz - 1.0) * X[3 * a + 2]};
        for (int i = 0; i < 3; ++i)
        {
            err = std::max(err, std::abs(u[3 * a + i] - ue[i]));
            umax = std::max(umax, std::abs(ue[i]));
        }
    }
    std::printf("      traction patch test: stretches %.12f %.12f, max |u - (F - I) X| / max|u| = %.3e\n", lz, l,
                err / umax);
    expect(err <= 1e-9 * umax, "uniaxial neo-Hooke under a top traction matches the closed form");
}
// the axial modes of a bar whose nodes move only along it, base fixed: the 1-D linear chain
// with consistent mass, lambda_k = 6 M_c / (rho h^2) (1 - cos t_k) / (2 + cos t_k),
// t_k = (2k - 1) pi / (2N), M_c = K + 4G/3
void bar_frequencies(neon_hip::context& ctx)
{
    int const N = 20;
    double const s = 0.1, bulk = 1.0 / (3 * (1 - 2 * 0.3)), shear = 1.0 / (2 * 1.3), rho = 1.0;
    std::int32_t nn = 0, ne = 0;
    neon_mesh_box_size(1, 1, N, s, s, double(N), 0.0, &nn, &ne);
    std::vector<double> X(3 * std::size_t(nn));
    std::vector<std::int32_t> conn(8 * std::size_t(ne)), ijk(3 * std::size_t(nn));
    neon_mesh_box(1, 1, N, s, s, double(N), 0.0, 1ull, 0.0, X.data(), conn.data(), ijk.data());
    std::vector<double> params(4 * std::size_t(ne), 0.0);
    for (std::int32_t e = 0; e < ne; ++e)
    {
        params[4 * e] = bulk;
        params[4 * e + 1] = shear;
    }
    neon_hip::device_mesh mesh(ctx, X, conn, params, NEON_MODEL_LINEAR_ELASTIC);
    std::vector<std::int32_t> fixed;
    for (std::int32_t a = 0; a < nn; ++a)
    {
        fixed.push_back(3 * a);
        fixed.push_back(3 * a + 1);
        if (ijk[3 * a + 2] == 0) fixed.push_back(3 * a + 2);
    }
    neon_hip::eigen_solver_options eo;
    eo.eigenvalues = 6;
    neon_hip::natural_frequency_matrix fm(mesh, eo, fixed, {rho});
    fm.solve();
    double err = 0.0;
    for (int k = 1; k <= 6; ++k)
    {
        double const t = (2 * k - 1) * M_PI / (2 * N);
        double const exact = 6 * (bulk + 4 * shear / 3) / rho * (1 - std::cos(t)) / (2 + std::cos(t));
        err = std::max(err, std::abs(fm.eigenvalues()[k - 1] - exact) / exact);
    }
    std::printf("      bar frequencies: %d inner solves, max relative error %.3e\n", fm.inner_solves(), err);
    expect(err < 1e-8, "natural frequencies of the bar match the 1-D consistent-mass chain");
}

void eigen_factory_errors()
{
    auto throws = [](neon_hip::eigen_solver_options const& o, char const* what) {
        try
        {
            neon_hip::check_eigen_solver(o);
        }
        catch (std::domain_error const& e)
        {
            return std::string(e.what()).find(what) != std::string::npos;
        }
        return false;
    };
    neon_hip::eigen_solver_options o;
    neon_hip::check_eigen_solver(o);
    o.eigenvalues = 0;
    expect(throws(o, "must be positive"), "zero eigenvalues are rejected");
    o.eigenvalues = 3;
    o.spectrum = "middle";
    expect(throws(o, "\"lower\" and \"upper\""), "an unknown spectrum is rejected");
    o.spectrum = "upper";
    o.type = "lanczos";
    expect(throws(o, "OpenCL"), "lanczos needs neon's OpenCL build");
}
// constitutive_model traits (host only): the flags and variable names of the reference's
// constructors (compressible_neohooke.hpp:39, small_strain_J2_plasticity.cpp:19-23,
// isotropic_linear_elasticity.cpp:17-20)
void model_traits()
{
    std::int32_t fd = -1, sym = -1;
    expect(neon_model_traits(NEON_MODEL_NEOHOOKE, &fd, &sym) == NEON_OK && fd == 1 && sym == 1,
           "neo-Hooke: finite deformation, symmetric");
    expect(neon_model_traits(NEON_MODEL_J2, &fd, &sym) == NEON_OK && fd == 0 && sym == 1,
           "small-strain J2: small deformation, symmetric");
    expect(std::string(neon_model_variable_names(NEON_MODEL_J2))
               == "effective_plastic_strain;linearised_plastic_strain;linearised_strain;von_mises_stress",
           "small-strain J2 variable names");
    expect(std::string(neon_model_variable_names(NEON_MODEL_NEOHOOKE)).empty(), "neo-Hooke adds no variable names");
    expect(neon_model_traits(99, &fd, &sym) == NEON_BAD_ARGUMENT && neon_model_variable_names(99) == nullptr,
           "unknown model refused");
}

// the constitutive_model mirror on a device mesh: traits, material, and its update through the
// C ABI (a homogeneous strain gives sigma = lambda tr(eps) I + 2 G eps at every point)
void constitutive_interface(neon_hip::context& ctx)
{
    std::int32_t nn = 0, ne = 0;
    neon_mesh_box_size(2, 2, 2, 1.0, 1.0, 1.0, 0.0, &nn, &ne);
    std::vector<double> X(3 * std::size_t(nn));
    std::vector<std::int32_t> conn(8 * std::size_t(ne)), ijk(3 * std::size_t(nn));
    neon_mesh_box(2, 2, 2, 1.0, 1.0, 1.0, 0.0, 1ull, 0.0, X.data(), conn.data(), ijk.data());
    double const K = 166.0e3, G = 80.0e3;
    std::vector<double> params;
    for (std::int32_t e = 0; e < ne; ++e) params.insert(params.end(), {K, G, 1.0e9, 0.0});
    neon_hip::device_mesh mesh(ctx, X, conn, params, NEON_MODEL_J2, NEON_TANGENT_REFERENCE, "steel");
    neon_hip::constitutive_model& cm = mesh.constitutive();
    expect(!cm.is_finite_deformation() && cm.is_symmetric() && mesh.is_symmetric(), "J2 mesh traits");
    expect(cm.intrinsic_material().name() == "steel" && cm.intrinsic_material().shear_modulus(3) == G,
           "intrinsic material");
    expect(cm.variable_names().count("effective_plastic_strain") == 1 && cm.variable_names().size() == 4,
           "variable names");
    double const ez = 1e-4;
    std::vector<double> u(3 * std::size_t(nn), 0.0);
    for (std::int32_t a = 0; a < nn; ++a) u[3 * a + 2] = ez * X[3 * a + 2];
    mesh.set_displacement(u);
    cm.update_internal_variables(0.0);
    std::vector<double> sigma(9 * 8 * std::size_t(ne));
    neon_hip::check(ctx.get(), neon_mesh_qp_state(ctx.get(), sigma.data(), nullptr, nullptr, nullptr, nullptr, nullptr));
    double const lam = K - 2.0 * G / 3.0, szz = (lam + 2.0 * G) * ez, sxx = lam * ez;
    double err = 0.0;
    for (std::size_t q = 0; q < 8 * std::size_t(ne); ++q)
        err = std::max({err, std::abs(sigma[9 * q + 8] - szz), std::abs(sigma[9 * q] - sxx), std::abs(sigma[9 * q + 1])});
    expect(err <= 1e-9 * szz, "update through constitutive_model gives C : eps");
}
} // namespace

int main()
{
    factory_errors();
    model_traits();
    eigen_factory_errors();
    load_vectors();
    int devices = 0;
    neon_device_count(&devices);
    if (devices == 0)
    {
        bool loud = false;
        try
        {
            neon_hip::context c(0);
        }
        catch (neon_hip::hip_error const& e)
        {
            loud = e.code == NEON_NO_DEVICE;
        }
        expect(loud, "context creation without a device throws hip_error(NO_DEVICE)");
        std::printf("no HIP device: device checks not run\n");
        return failures ? 1 : 77;
    }
    neon_hip::context ctx(0);
    tridiagonal(ctx);
    constitutive_interface(ctx);
    patch_test(ctx);
    traction_patch_test(ctx);
    bar_frequencies(ctx);
    std::printf("%d failure(s)\n", failures);
    return failures ? 1 : 0;
}
