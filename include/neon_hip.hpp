/*
 * neon_hip.hpp — C++17 host classes over the C ABI of neon_hip.h, shaped like dbeurle/neon's
 * own interfaces so that code written against those switches by type name (paths relative to
 * the reference repository root):
 *
 *   linear_solver, iterative_linear_solver  src/solver/linear/linear_solver.hpp:13-56
 *   conjugate_gradient_hip                  conjugate_gradient_cuda (conjugate_gradient_cuda.hpp:17-27)
 *   biconjugate_gradient_stabilised_hip     biconjugate_gradient_stabilised_cuda
 *                                           (biconjugate_gradient_stabilised_cuda.hpp)
 *   make_linear_solver                      src/solver/linear/linear_solver_factory.cpp:28-174
 *   computational_error, hip_error          computational_error, cuda_error (src/exceptions.hpp:14-23)
 *   device_mesh                             mechanics::solid::mesh + its constitutive models
 *                                           (src/mesh/mechanics/solid/mesh.hpp:37-80)
 *   constitutive_model, material_property   src/constitutive/constitutive_model.hpp:19-56,
 *                                           src/material/material_property.hpp:13-37
 *   static_matrix                           fem::static_matrix (src/assembler/mechanics/static_matrix.hpp:145-467)
 *
 * Recoverable failures (Krylov or Newton limits, non-finite values, det F < 0, radial return)
 * throw computational_error, which neon's load-step driver catches to cut back; everything else
 * throws hip_error (a std::domain_error, like neon's cuda_error).  Host containers are borrowed
 * for the duration of a call; the context owns every device buffer.
 */
#pragma once

#include "neon_hip.h"

#include <cstdint>
#include <functional>
#include <memory>
#include <optional>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace neon_hip
{
/// A failure that a smaller load increment may cure (neon's computational_error).
struct computational_error : public std::domain_error
{
    computational_error(int c, std::string const& what) : std::domain_error(what), code(c) {}
    int code;
};

/// Any other failure: bad arguments, HIP runtime errors, exhausted cut-backs (neon's cuda_error).
struct hip_error : public std::domain_error
{
    hip_error(int c, std::string const& what) : std::domain_error(what), code(c) {}
    int code;
};

inline void check(neon_context* ctx, int code)
{
    if (code == NEON_OK) return;
    std::string const msg = ctx ? neon_context_last_error(ctx) : (code == NEON_NO_DEVICE ? "no HIP device" : "");
    if (code > 0) throw computational_error(code, msg);
    throw hip_error(code, msg);
}

/// One device and one HIP stream; owns every device buffer made through it.
class context
{
public:
    explicit context(int device = 0) { check(nullptr, neon_context_create(device, &ctx_)); }
    ~context()
    {
        if (ctx_) neon_context_destroy(ctx_);
    }
    context(context const&) = delete;
    context& operator=(context const&) = delete;
    neon_context* get() const { return ctx_; }

    /// The process-wide context on device 0, created on first use (the CUDA solvers find
    /// their device the same way, conjugate_gradient_cuda.cpp:15-51).
    static context& default_context()
    {
        static context c(0);
        return c;
    }

private:
    neon_context* ctx_ = nullptr;
};

/// Row-major CSR with int32 indices and sorted columns: the layout of neon's sparse_matrix
/// (Eigen::SparseMatrix<double, RowMajor>, src/numeric/sparse_matrix.hpp:15), borrowed.
struct csr_view
{
    std::int32_t rows = 0;
    std::int64_t nnz = 0;
    const std::int32_t* row_ptr = nullptr;
    const std::int32_t* col = nullptr;
    const double* val = nullptr;
};

class linear_solver
{
public:
    virtual ~linear_solver() = default;

    virtual void solve(csr_view const& A, std::vector<double>& x, std::vector<double> const& b) = 0;

    /// Notifies the solver of a change in the sparsity structure of A
    void update_sparsity_pattern() { build_sparsity_pattern = true; }

protected:
    bool build_sparsity_pattern{true};
};

class iterative_linear_solver : public linear_solver
{
public:
    explicit iterative_linear_solver() = default;
    explicit iterative_linear_solver(double const tolerance) : residual_tolerance(tolerance) {}
    explicit iterative_linear_solver(std::int32_t const iterations) : max_iterations(iterations) {}
    explicit iterative_linear_solver(double const tolerance, std::int32_t const iterations)
        : residual_tolerance(tolerance), max_iterations(iterations)
    {
    }

    /// NEON_STOP_ABSOLUTE_PRECONDITIONED (the CUDA solvers' rule, the default) or
    /// NEON_STOP_RELATIVE (the CPU solvers' rule)
    int criterion = NEON_STOP_ABSOLUTE_PRECONDITIONED;
    /// NEON_PRECISION_FP64 or NEON_PRECISION_MIXED (conjugate gradient only)
    int precision = NEON_PRECISION_FP64;
    /// device context; the process-wide one when null
    context* ctx = nullptr;

    /// Krylov iterations and final residual measure of the last solve
    std::int32_t iterations() const { return iterations_; }
    double error() const { return error_; }

protected:
    void run(int solver, csr_view const& A, std::vector<double>& x, std::vector<double> const& b)
    {
        if (b.size() != std::size_t(A.rows)) throw hip_error(NEON_BAD_ARGUMENT, "right-hand side size mismatch");
        context& c = ctx ? *ctx : context::default_context();
        x.assign(A.rows, 0.0);
        check(c.get(), neon_krylov_solve_csr(c.get(), solver, criterion, precision, A.rows, A.nnz, A.row_ptr, A.col,
                                             A.val, b.data(), x.data(), residual_tolerance, max_iterations,
                                             &iterations_, &error_));
        build_sparsity_pattern = false;
    }

    double residual_tolerance{1.0e-5};
    std::int32_t max_iterations{2000};
    std::int32_t iterations_ = 0;
    double error_ = 0.0;
};

/// Jacobi-preconditioned conjugate gradient on the GPU (drop-in for conjugate_gradient_cuda).
class conjugate_gradient_hip : public iterative_linear_solver
{
public:
    using iterative_linear_solver::iterative_linear_solver;

    void solve(csr_view const& A, std::vector<double>& x, std::vector<double> const& b) override
    {
        run(NEON_SOLVER_CG, A, x, b);
    }
};

/// Jacobi-preconditioned BiCGSTAB on the GPU (drop-in for biconjugate_gradient_stabilised_cuda).
class biconjugate_gradient_stabilised_hip : public iterative_linear_solver
{
public:
    using iterative_linear_solver::iterative_linear_solver;

    void solve(csr_view const& A, std::vector<double>& x, std::vector<double> const& b) override
    {
        run(NEON_SOLVER_BICGSTAB, A, x, b);
    }
};

/// The fields of neon's "linear_solver" JSON object that select a GPU solver
/// ({"type": "iterative", "device": "gpu", "backend": {"type": "hip"}, ...}).
struct solver_options
{
    std::string type;
    std::string device;
    std::string backend;
    std::string precision = "fp64";
    std::optional<double> tolerance;
    std::optional<std::int32_t> maximum_iterations;
};

namespace detail
{
template <typename Solver>
std::unique_ptr<iterative_linear_solver> make_with(solver_options const& o)
{
    // make_iterative_solver (linear_solver_factory.cpp:28-65): the constructor matching the
    // options given
    if (o.tolerance && o.maximum_iterations) return std::make_unique<Solver>(*o.tolerance, *o.maximum_iterations);
    if (o.tolerance) return std::make_unique<Solver>(*o.tolerance);
    if (o.maximum_iterations) return std::make_unique<Solver>(*o.maximum_iterations);
    return std::make_unique<Solver>();
}
} // namespace detail

/// make_linear_solver (linear_solver_factory.cpp:67-174) for the GPU iterative solvers: conjugate
/// gradient when the system is symmetric, BiCGSTAB otherwise.
inline std::unique_ptr<linear_solver> make_linear_solver(solver_options const& o, bool const is_symmetric)
{
    if (o.type.empty())
        throw std::domain_error("Did not find a linear solver type.  Did you try specifying \"type\" for the "
                                "linear solver name?");
    if (o.type != "iterative" || o.device != "gpu")
        throw std::domain_error("this library provides the iterative gpu solvers only");
    if (o.backend.empty()) throw std::domain_error("\"backend\" requires a \"type\" option");
    if (o.backend != "hip") throw std::domain_error("Incorrect backend option specified.");
    int precision;
    if (o.precision == "fp64")
        precision = NEON_PRECISION_FP64;
    else if (o.precision == "mixed")
        precision = NEON_PRECISION_MIXED;
    else
        throw std::domain_error("\"precision\" must be \"fp64\" or \"mixed\"");
    if (precision == NEON_PRECISION_MIXED && !is_symmetric)
        throw std::domain_error("mixed precision supports the conjugate gradient only");
    std::unique_ptr<iterative_linear_solver> s;
    if (is_symmetric)
        s = detail::make_with<conjugate_gradient_hip>(o);
    else
        s = detail::make_with<biconjugate_gradient_stabilised_hip>(o);
    s->precision = precision;
    return s;
}

/// The parameters of a device mesh's material (material_property and its isotropic subclasses,
/// src/material/material_property.hpp:13-37, isotropic_elastic_property.hpp): the name and, per
/// element (the reference holds one value per part), the model's four parameters in the order of
/// neon_mesh_create.
class material_property
{
public:
    material_property(std::string name, std::vector<double> params, std::optional<double> density = std::nullopt)
        : name_(std::move(name)), params_(std::move(params)), density_(density)
    {
    }
    std::string const& name() const { return name_; }
    double initial_density() const
    {
        if (!density_) throw std::domain_error("Density was requested, but not specified in the input file");
        return *density_;
    }
    /// parameter k (0..3) of element e
    double parameter(std::int32_t e, int k) const { return params_[4 * std::size_t(e) + k]; }
    double bulk_modulus(std::int32_t e) const { return parameter(e, 0); }
    double shear_modulus(std::int32_t e) const { return parameter(e, 1); }
    double lambda(std::int32_t e) const { return parameter(e, 0) - 2.0 * parameter(e, 1) / 3.0; }
    std::vector<double> const& parameters() const { return params_; }

private:
    std::string name_;
    std::vector<double> params_;
    std::optional<double> density_;
};

class device_mesh;

/// constitutive_model (src/constitutive/constitutive_model.hpp:19-56) of a device mesh: its
/// update launches the device kernels on every quadrature point of the mesh, whose internal
/// variables stay on the device.
class constitutive_model
{
public:
    constitutive_model(device_mesh& m, int model, material_property material);

    /// Update the internal variables and tangents at the quadrature points for the displacement
    /// the mesh holds (time_step_size: the rate models' step; neon passes the load increment at a
    /// step's first update and 0 otherwise)
    void update_internal_variables(double const time_step_size);

    material_property const& intrinsic_material() const { return material_; }
    bool is_finite_deformation() const { return finite_; }
    bool is_symmetric() const { return symmetric_; }
    std::set<std::string> const& variable_names() const noexcept { return names_; }

private:
    device_mesh& mesh_;
    material_property material_;
    bool finite_ = false, symmetric_ = true;
    std::set<std::string> names_;
};

/// Device-resident hex8 mesh with its constitutive model, sparsity pattern and tangent.
class device_mesh
{
public:
    /// X: 3 per node; conn: 8 per element (hexahedron.cpp:21-28 order); params: 4 per element
    /// (bulk, shear, yield stress, hardening modulus); model: NEON_MODEL_*; tangent: NEON_TANGENT_*
    device_mesh(context& c, std::vector<double> const& X, std::vector<std::int32_t> const& conn,
                std::vector<double> const& params, int model, int tangent = NEON_TANGENT_REFERENCE,
                std::string material_name = "")
        : ctx_(c), nodes_(std::int32_t(X.size() / 3)), elements_(std::int32_t(conn.size() / 8))
    {
        if (X.size() % 3 || conn.size() % 8 || params.size() != 4 * conn.size() / 8)
            throw hip_error(NEON_BAD_ARGUMENT, "mesh array sizes disagree");
        check(c.get(), neon_mesh_create(c.get(), nodes_, X.data(), elements_, conn.data(), model, params.data(),
                                        tangent));
        model_ = std::make_unique<constitutive_model>(*this, model, material_property(std::move(material_name), params));
    }

    /// the mesh's constitutive model (submesh::constitutive, submesh.hpp:62)
    constitutive_model& constitutive() { return *model_; }
    constitutive_model const& constitutive() const { return *model_; }
    /// mesh::is_symmetric (mesh.cpp:47-53): every submesh's model is symmetric (one model per
    /// device mesh here)
    bool is_symmetric() const { return model_->is_symmetric(); }

    std::int32_t nodes() const { return nodes_; }
    std::int32_t elements() const { return elements_; }
    std::int64_t dofs() const { return 3ll * nodes_; }
    context& ctx() const { return ctx_; }

    void set_parameters(std::vector<double> const& params)
    {
        check(ctx_.get(), neon_mesh_set_parameters(ctx_.get(), params.data()));
    }
    /// x = X + u (material_coordinates::update_current_configuration, material_coordinates.cpp:23-27)
    void set_displacement(std::vector<double> const& u)
    {
        check(ctx_.get(), neon_mesh_set_displacement(ctx_.get(), u.data()));
    }
    /// x = X + u and the constitutive update at every quadrature point (mesh.cpp:57-71)
    void update_internal_variables(std::vector<double> const& u, double const time_step = 0.0)
    {
        check(ctx_.get(), neon_mesh_set_displacement(ctx_.get(), u.data()));
        check(ctx_.get(), neon_mesh_update_internal_variables(ctx_.get(), time_step));
    }
    /// commit (converged) or revert the history (mesh.cpp:73-77)
    void save_internal_variables(bool const have_converged)
    {
        check(ctx_.get(), neon_mesh_save_internal_variables(ctx_.get(), have_converged ? 1 : 0));
    }
    /// the ageing microsphere's scission and recombination probabilities
    /// (micromechanical_elastomer.cpp:22-53)
    void set_ageing(double const scission, double const recombination)
    {
        check(ctx_.get(), neon_mesh_set_ageing(ctx_.get(), scission, recombination));
    }
    /// the ageing microsphere's network state, 11 values per quadrature point
    std::vector<double> ageing_state() const
    {
        std::vector<double> z(88 * std::size_t(elements_));
        check(ctx_.get(), neon_mesh_ageing_state(ctx_.get(), z.data()));
        return z;
    }
    /// gradient of an objective with respect to the per-element parameters (4 per element) at
    /// the held state, through the adjoint of the consistent tangent
    std::vector<double> adjoint_gradient(std::vector<double> const& dJdu, std::vector<std::int32_t> const& fixed,
                                         double const tolerance = 1e-12, std::int32_t const max_iterations = 20000)
    {
        std::vector<double> grad(4 * std::size_t(elements_));
        std::int32_t its = 0;
        check(ctx_.get(), neon_adjoint_gradient(ctx_.get(), dJdu.data(), std::int32_t(fixed.size()), fixed.data(),
                                                tolerance, max_iterations, grad.data(), &its));
        return grad;
    }

private:
    context& ctx_;
    std::int32_t nodes_, elements_;
    std::unique_ptr<constitutive_model> model_;
};

inline constitutive_model::constitutive_model(device_mesh& m, int model, material_property material)
    : mesh_(m), material_(std::move(material))
{
    std::int32_t fd = 0, sym = 1;
    if (neon_model_traits(model, &fd, &sym) != NEON_OK) throw hip_error(NEON_BAD_ARGUMENT, "unknown constitutive model");
    finite_ = fd != 0;
    symmetric_ = sym != 0;
    std::string const all = neon_model_variable_names(model);
    for (std::size_t b = 0; b < all.size();)
    {
        std::size_t const e = std::min(all.find(';', b), all.size());
        names_.insert(all.substr(b, e - b));
        b = e + 1;
    }
}

inline void constitutive_model::update_internal_variables(double const time_step_size)
{
    check(mesh_.ctx().get(), neon_mesh_update_internal_variables(mesh_.ctx().get(), time_step_size));
}

/// A prescribed displacement history on a set of dofs (boundary.cpp:50-85).
struct dirichlet_boundary
{
    std::vector<std::int32_t> dofs;
    std::vector<double> times, values;
};

/// A nonfollower load (neon's traction, pressure, body_force or nodal_force boundary,
/// nonfollower_load.cpp:12-122): f_ext += load(t) * values at dofs, where values is the load
/// vector of a unit load (surface_load_vector / volume_load_vector, or 1 per dof for a nodal
/// force) and load(t) interpolates (times, loads).  fixed_time >= 0 evaluates the history at
/// that time in every step (neon applies nodal forces at t = 1, static_matrix.hpp:230-236).
struct nonfollower_load
{
    std::vector<std::int32_t> dofs;
    std::vector<double> values, times, loads;
    double fixed_time = -1.0;
};

/// Unit-load vector (3N, dense) of quadrilateral faces on the initial configuration X:
/// traction in component 0..2, or unit pressure for component -1 (neon_surface_load_vector).
inline std::vector<double> surface_load_vector(std::vector<double> const& X, std::vector<std::int32_t> const& faces,
                                               std::int32_t component, bool reduced = false)
{
    std::vector<double> out(X.size(), 0.0);
    if (neon_surface_load_vector(std::int32_t(X.size() / 3), X.data(), std::int32_t(faces.size() / 4), faces.data(),
                                 component, reduced ? 1 : 0, out.data())
        != NEON_OK)
        throw std::domain_error("invalid surface load (node index or component out of range)");
    return out;
}

/// Unit body-force vector (3N) of hex8 elements (neon_volume_load_vector).
inline std::vector<double> volume_load_vector(std::vector<double> const& X, std::vector<std::int32_t> const& conn,
                                              std::int32_t component, bool reduced = false)
{
    std::vector<double> out(X.size(), 0.0);
    if (neon_volume_load_vector(std::int32_t(X.size() / 3), X.data(), std::int32_t(conn.size() / 8), conn.data(),
                                component, reduced ? 1 : 0, out.data())
        != NEON_OK)
        throw std::domain_error("invalid body force (node index or component out of range)");
    return out;
}

/// The "time", "nonlinear_options" and "linear_solver" entries of a neon simulation.
struct simulation_options
{
    double period = 1.0, initial = 1.0, minimum = 1.0, maximum = 1.0;
    bool adaptive = false;
    double residual_tolerance = 1e-3, displacement_tolerance = 1e-3;
    std::int32_t newton_iterations = 10;
    /// neon's "absolute_tolerance" key: absolute norms (static_matrix.hpp:385-400)
    bool absolute_tolerance = false;
    /// extension: relative norms without neon's fall-backs to absolute ones
    bool always_relative = false;
    /// extension: keep the load history on the device for the J2 adjoint
    bool record_trajectory = false;
    solver_options linear;
    /// "latin_search_direction" of a LATIN step (latin_matrix.hpp:95-104)
    std::optional<double> latin_search_direction;
};

/// Newton-Raphson with adaptive load stepping on the device (fem::static_matrix::solve).
class static_matrix
{
public:
    static_matrix(device_mesh& m, simulation_options const& o, std::vector<dirichlet_boundary> b,
                  std::vector<nonfollower_load> loads = {})
        : mesh_(m), options_(o), boundaries_(std::move(b)), loads_(std::move(loads))
    {
    }

    /// Solves every load step; returns the converged displacement.  A recoverable failure that
    /// survives the cut-backs throws hip_error (neon's "Too many reductions" domain_error).
    std::vector<double> solve(std::int32_t const record_capacity = 4096)
    {
        solver_options const& l = options_.linear;
        if (l.type != "iterative" || l.device != "gpu" || l.backend != "hip")
            throw std::domain_error("the device Newton driver needs the iterative hip gpu solver");
        neon_newton_options o{};
        o.period = options_.period;
        o.initial = options_.initial;
        o.minimum = options_.minimum;
        o.maximum = options_.maximum;
        o.adaptive = options_.adaptive ? 1 : 0;
        o.residual_tolerance = options_.residual_tolerance;
        o.displacement_tolerance = options_.displacement_tolerance;
        o.max_newton_iterations = options_.newton_iterations;
        o.absolute_norm = options_.absolute_tolerance ? 1 : 0;
        o.norm_mode = options_.always_relative ? 1 : 0;
        o.solver_tolerance = l.tolerance.value_or(1.0e-5);
        o.solver_max_iterations = l.maximum_iterations.value_or(2000);
        o.solver_criterion = NEON_STOP_ABSOLUTE_PRECONDITIONED;
        // make_linear_solver(linear_solver, mesh.is_symmetric()) (static_matrix.hpp:109-112,
        // linear_solver_factory.cpp:150-161)
        o.solver = mesh_.is_symmetric() ? NEON_SOLVER_CG : NEON_SOLVER_BICGSTAB;
        o.precision = l.precision == "mixed" ? NEON_PRECISION_MIXED : NEON_PRECISION_FP64;
        if (o.precision == NEON_PRECISION_MIXED) o.solver_criterion = NEON_STOP_RELATIVE;
        o.record_trajectory = options_.record_trajectory ? 1 : 0;
        if (on_step)
        {
            // neon calls mesh.write(step, time) here (static_matrix.hpp:159-186)
            o.on_step = [](void* user, std::int32_t step, double time) {
                (*static_cast<std::function<void(std::int32_t, double)>*>(user))(step, time);
            };
            o.callback_user = &on_step;
        }
        std::vector<neon_dirichlet> bcs;
        for (auto const& b : boundaries_)
            bcs.push_back(neon_dirichlet{std::int32_t(b.dofs.size()), b.dofs.data(), std::int32_t(b.times.size()),
                                         b.times.data(), b.values.data()});
        std::vector<neon_neumann> lds;
        for (auto const& f : loads_)
        {
            if (f.dofs.size() != f.values.size() || f.times.size() != f.loads.size())
                throw std::domain_error("nonfollower load: dofs/values or times/loads differ in length");
            lds.push_back(neon_neumann{std::int32_t(f.dofs.size()), f.dofs.data(), f.values.data(),
                                       std::int32_t(f.times.size()), f.times.data(), f.loads.data(), f.fixed_time});
        }
        std::vector<double> u(mesh_.dofs());
        records_.assign(record_capacity, neon_newton_record{});
        std::int32_t n = 0;
        int const code =
            options_.latin_search_direction
                ? neon_latin_solve(mesh_.ctx().get(), &o, *options_.latin_search_direction, std::int32_t(bcs.size()),
                                   bcs.data(), std::int32_t(lds.size()), lds.data(), u.data(), records_.data(),
                                   record_capacity, &n, &timing_)
                : neon_newton_solve_loaded(mesh_.ctx().get(), &o, std::int32_t(bcs.size()), bcs.data(),
                                           std::int32_t(lds.size()), lds.data(), u.data(), records_.data(),
                                           record_capacity, &n, &timing_);
        records_.resize(std::min(n, record_capacity));
        check(mesh_.ctx().get(), code == NEON_OK || code < 0 ? code : NEON_FATAL);
        return u;
    }

    /// called with (step, time) where neon writes its output: the initial state, after a
    /// cut-back and after each converged step; the device state (neon_mesh_displacement,
    /// neon_mesh_qp_state, neon_mesh_download) is current during the call
    std::function<void(std::int32_t, double)> on_step;

    /// one record per Newton iteration (load step, iteration, Krylov count, norms)
    std::vector<neon_newton_record> const& records() const { return records_; }
    neon_newton_timing const& timing() const { return timing_; }

private:
    device_mesh& mesh_;
    simulation_options options_;
    std::vector<dirichlet_boundary> boundaries_;
    std::vector<nonfollower_load> loads_;
    std::vector<neon_newton_record> records_;
    neon_newton_timing timing_{};
};
/// LATIN iterations on the device (latin_matrix.hpp:150-244): static_matrix's load stepping
/// with the LATIN force as the right-hand side after each step's first iteration.
class latin_matrix : public static_matrix
{
public:
    latin_matrix(device_mesh& m, simulation_options const& o, std::vector<dirichlet_boundary> b,
                 std::vector<nonfollower_load> loads = {})
        : static_matrix(m, checked(o), std::move(b), std::move(loads))
    {
    }

private:
    static simulation_options const& checked(simulation_options const& o)
    {
        if (!o.latin_search_direction)
            throw std::domain_error("latin_search_direction not specified in nonlinear_options");
        return o;
    }
};

/// The "eigen_solver" block of a natural-frequency step (eigen_solver_factory.cpp:12-77).
struct eigen_solver_options
{
    std::string type = "arpack";
    std::int64_t eigenvalues = 10;
    std::string spectrum = "lower";
};

/// make_eigen_solver's checks: "arpack" maps to the device solver; "lanczos" and
/// "power_iteration" need neon's OpenCL build and solve no generalised problem.
inline void check_eigen_solver(eigen_solver_options const& o)
{
    if (o.eigenvalues <= 0) throw std::domain_error("The number of requested eigenvalues must be positive");
    if (o.spectrum != "lower" && o.spectrum != "upper")
        throw std::domain_error("A spectrum was requested but only \"lower\" and \"upper\" options are supported");
    if (o.type == "power_iteration" || o.type == "lanczos")
        throw std::domain_error("The \"" + o.type + "\" option requires OpenCL support.  Recompile with -DENABLE_OPENCL=1");
    if (o.type.empty())
        throw std::domain_error("Eigen solver type was not provided.  Please use \"power_iteration\", \"arpack\" or "
                                "\"lanczos\"");
    if (o.type != "arpack")
        throw std::domain_error("Eigen solver type \"" + o.type + "\" is not recognised.  Please use \"arpack\"");
}

/// fem::natural_frequency_matrix (natural_frequency_matrix.hpp:17-161) on a device mesh: K at the
/// mesh's held state and the consistent mass with neon's homogeneous Dirichlet treatment, then the
/// eigenvalues of K x = lambda M x (increasing) and M-orthonormal mode shapes (3N per mode).
class natural_frequency_matrix
{
public:
    natural_frequency_matrix(device_mesh& m, eigen_solver_options o, std::vector<std::int32_t> fixed,
                             std::vector<double> density)
        : mesh_(m), options_(std::move(o)), fixed_(std::move(fixed)), density_(std::move(density))
    {
        check_eigen_solver(options_);
        if (density_.size() == 1) density_.assign(std::size_t(m.elements()), density_[0]);
        if (density_.size() != std::size_t(m.elements())) throw hip_error(NEON_BAD_ARGUMENT, "one density per element");
    }

    void solve(double const tolerance = 1.0e-8, std::int32_t const max_basis = 0, std::int32_t const block = 0)
    {
        std::int32_t const k = std::int32_t(options_.eigenvalues);
        eigenvalues_.assign(k, 0.0);
        eigenvectors_.assign(std::size_t(k) * mesh_.dofs(), 0.0);
        check(mesh_.ctx().get(),
              neon_natural_frequency(mesh_.ctx().get(), density_.data(), std::int32_t(fixed_.size()), fixed_.data(), k,
                                     options_.spectrum == "upper" ? NEON_SPECTRUM_UPPER : NEON_SPECTRUM_LOWER,
                                     tolerance, max_basis, block, eigenvalues_.data(), eigenvectors_.data(), &solves_));
    }

    std::vector<double> const& eigenvalues() const { return eigenvalues_; }
    /// mode e at [e * 3N, (e + 1) * 3N)
    std::vector<double> const& eigenvectors() const { return eigenvectors_; }
    std::int32_t inner_solves() const { return solves_; }

private:
    device_mesh& mesh_;
    eigen_solver_options options_;
    std::vector<std::int32_t> fixed_;
    std::vector<double> density_;
    std::vector<double> eigenvalues_, eigenvectors_;
    std::int32_t solves_ = 0;
};
} // namespace neon_hip
