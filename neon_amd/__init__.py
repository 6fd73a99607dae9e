"""neon_amd — MI355X (gfx950) implicit Newton step for hex8 solid mechanics.

Python mirror of the reference's host interfaces over the C ABI of ``include/neon_hip.h``:

* :class:`conjugate_gradient_hip` / :class:`biconjugate_gradient_stabilised_hip` — the
  ``linear_solver`` subclasses selected like ``conjugate_gradient_cuda``
  (src/solver/linear/linear_solver.hpp:13-56, src/solver/linear/conjugate_gradient_cuda.hpp:17-59),
  with :func:`make_linear_solver` following src/solver/linear/linear_solver_factory.cpp:67-174;
* :class:`DeviceMesh` — the device-resident ``mesh``/``submesh``/``constitutive_model`` state
  (src/mesh/mechanics/solid/mesh.hpp:37-80, submesh.hpp:38-67);
* :class:`static_matrix` — the Newton driver (src/assembler/mechanics/static_matrix.hpp:29-107).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import (  # noqa: F401
    MODEL_J2,
    MODEL_LINEAR_ELASTIC,
    MODEL_MICROSPHERE_GAUSSIAN,
    MODEL_MICROSPHERE_LANGEVIN,
    MODEL_MICROSPHERE_AGEING,
    MODEL_MICROSPHERE_NONAFFINE,
    MODEL_J2_DAMAGE,
    MODEL_NEOHOOKE,
    SPHERE_BO33,
    PRECISION_FP64,
    PRECISION_MIXED,
    SOLVER_BICGSTAB,
    SOLVER_CG,
    SPECTRUM_LOWER,
    SPECTRUM_UPPER,
    SPMV_SLICED,
    SPMV_INDEX32,
    SPMV_NONTEMPORAL,
    SPMV_SYMMETRIC,
    STOP_ABSOLUTE_PRECONDITIONED,
    STOP_RELATIVE,
    TANGENT_CONSISTENT,
    TANGENT_REFERENCE,
    NeonError,
)

__all__ = [
    "Context",
    "DeviceMesh",
    "conjugate_gradient_hip",
    "biconjugate_gradient_stabilised_hip",
    "make_linear_solver",
    "constitutive_model",
    "material_property",
    "model_traits",
    "static_matrix",
    "latin_matrix",
    "natural_frequency_matrix",
    "make_eigen_solver",
    "computational_error",
]


class computational_error(RuntimeError):
    """Recoverable numerical failure (neon's ``computational_error``, src/exceptions.hpp:14-17)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{_capi.STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


def _raise(code: int, ctx_handle) -> None:
    if code == _capi.NEON_OK:
        return
    msg = _capi.lib().neon_context_last_error(ctx_handle) if ctx_handle else b""
    msg = msg.decode() if msg else ""
    if code > 0:
        raise computational_error(code, msg)
    raise NeonError(code, msg)


class Context:
    """One device, one HIP stream (``neon_context_create``)."""

    def __init__(self, device: int = 0):
        self.handle = C.c_void_p()
        code = _capi.lib().neon_context_create(device, C.byref(self.handle))
        if code != _capi.NEON_OK:
            raise NeonError(code, "neon_context_create failed")

    def close(self):
        if self.handle:
            _capi.lib().neon_context_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, code: int) -> None:
        _raise(code, self.handle)

    def synchronize(self) -> None:
        self.check(_capi.lib().neon_context_synchronize(self.handle))

    def profile(self, enable: bool) -> None:
        """Time SpMV launches and assemblies with HIP events on the context's stream."""
        self.check(_capi.lib().neon_context_profile(self.handle, int(enable)))

    def profile_read(self, kernel_class: int, reset: bool = True):
        """(launches, summed milliseconds) for a kernel class (``_capi.KERNEL_*``)."""
        cnt, ms = C.c_int64(0), C.c_double(0.0)
        self.check(_capi.lib().neon_context_profile_read(self.handle, kernel_class, int(reset), C.byref(cnt),
                                                         C.byref(ms)))
        return cnt.value, ms.value


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _csr_arrays(A):
    """(row_ptr, col, val) from a scipy CSR matrix or a tuple."""
    if isinstance(A, tuple):
        rp, ci, v = A
    else:
        A = A.tocsr()
        rp, ci, v = A.indptr, A.indices, A.data
    rp = np.ascontiguousarray(rp, dtype=np.int32)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.float64)
    return rp, ci, v


class iterative_linear_solver:
    """``iterative_linear_solver`` (src/solver/linear/linear_solver.hpp:27-56): defaults
    tolerance 1e-5 and 2000 iterations."""

    solver_kind = SOLVER_CG

    def __init__(self, residual_tolerance: float = 1.0e-5, max_iterations: int = 2000,
                 criterion: int = STOP_ABSOLUTE_PRECONDITIONED, precision: int = PRECISION_FP64,
                 context: Context | None = None):
        self.residual_tolerance = float(residual_tolerance)
        self.max_iterations = int(max_iterations)
        self.criterion = criterion
        self.precision = precision
        self.ctx = context or default_context()
        self.iterations = 0
        self.error = 0.0

    def solve(self, A, x: np.ndarray, b: np.ndarray) -> None:
        """``solve(sparse_matrix const& A, vector& x, vector const& b)``: x is overwritten.
        Raises :class:`computational_error` where neon throws it."""
        rp, ci, v = _csr_arrays(A)
        n = len(rp) - 1
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.zeros(n)
        its = C.c_int32(0)
        err = C.c_double(0.0)
        code = _capi.lib().neon_krylov_solve_csr(
            self.ctx.handle, self.solver_kind, self.criterion, self.precision, n, int(rp[-1]),
            _capi.i32(rp), _capi.i32(ci), _capi.f64(v), _capi.f64(b), _capi.f64(out),
            self.residual_tolerance, self.max_iterations, C.byref(its), C.byref(err))
        self.iterations, self.error = its.value, err.value
        x[...] = out
        self.ctx.check(code)


class conjugate_gradient_hip(iterative_linear_solver):
    """Drop-in for ``conjugate_gradient_cuda`` (src/solver/linear/conjugate_gradient_cuda.cpp:72-192)."""

    solver_kind = SOLVER_CG


class biconjugate_gradient_stabilised_hip(iterative_linear_solver):
    """Drop-in for ``biconjugate_gradient_stabilised_cuda``
    (src/solver/linear/biconjugate_gradient_stabilised_cuda.cpp:31-200)."""

    solver_kind = SOLVER_BICGSTAB


def make_linear_solver(solver_data: dict, is_symmetric: bool, context: Context | None = None):
    """The ``"iterative"`` / ``"device": "gpu"`` branch of make_linear_solver
    (src/solver/linear/linear_solver_factory.cpp:67-174) with ``backend.type == "hip"``;
    ``backend.precision`` selects ``"fp64"`` (default) or ``"mixed"``."""
    if "type" not in solver_data:
        raise ValueError('A "linear_solver" "type" must be specified')
    if solver_data["type"] != "iterative":
        raise ValueError("only the iterative solvers run on the HIP backend")
    if solver_data.get("device", "cpu") != "gpu":
        raise ValueError('the HIP solvers need "device": "gpu"')
    backend = solver_data.get("backend")
    if backend is None:
        raise ValueError('The gpu linear solvers require a "backend" object')
    if backend.get("type") != "hip":
        raise ValueError("Incorrect backend option specified.")
    precision = {"fp64": PRECISION_FP64, "mixed": PRECISION_MIXED}[backend.get("precision", "fp64")]
    kwargs = {}
    if "tolerance" in solver_data:
        kwargs["residual_tolerance"] = solver_data["tolerance"]
    if "maximum_iterations" in solver_data:
        kwargs["max_iterations"] = solver_data["maximum_iterations"]
    cls = conjugate_gradient_hip if is_symmetric else biconjugate_gradient_stabilised_hip
    return cls(precision=precision, context=context, **kwargs)


def model_traits(model: int) -> tuple[bool, bool]:
    """(is_finite_deformation, is_symmetric) of a constitutive model (``neon_model_traits``)."""
    fd, sym = C.c_int32(0), C.c_int32(0)
    code = _capi.lib().neon_model_traits(int(model), C.byref(fd), C.byref(sym))
    if code != 0:
        raise NeonError(code, f"unknown constitutive model {model}")
    return bool(fd.value), bool(sym.value)


class material_property:
    """The parameters of a device mesh's material (``material_property`` and its isotropic
    subclasses, src/material/material_property.hpp:13-37, isotropic_elastic_property.hpp):
    ``name`` and, per element (the reference holds one value per part), the model's four
    parameters in the order of ``neon_mesh_create``."""

    def __init__(self, name: str, params: np.ndarray, density: float | None = None):
        self._name = name
        self.params = np.asarray(params).reshape(-1, 4)
        self._density = density

    def name(self) -> str:
        return self._name

    def initial_density(self) -> float:
        if self._density is None:
            raise ValueError("Density was requested, but not specified in the input file")
        return self._density

    def bulk_modulus(self) -> np.ndarray:
        return self.params[:, 0]

    def shear_modulus(self) -> np.ndarray:
        return self.params[:, 1]

    def lambda_(self) -> np.ndarray:
        return self.params[:, 0] - 2.0 * self.params[:, 1] / 3.0

    def elastic_modulus(self) -> np.ndarray:
        K, G = self.params[:, 0], self.params[:, 1]
        return 9.0 * K * G / (3.0 * K + G)

    def poissons_ratio(self) -> np.ndarray:
        K, G = self.params[:, 0], self.params[:, 1]
        return (3.0 * K - 2.0 * G) / (2.0 * (3.0 * K + G))


class constitutive_model:
    """``constitutive_model`` (src/constitutive/constitutive_model.hpp:19-56) of a
    :class:`DeviceMesh`: its update launches the device kernels on the mesh's quadrature points,
    whose state lives on the device (``DeviceMesh.qp_state`` reads it)."""

    def __init__(self, mesh: "DeviceMesh", material_name: str = "", density: float | None = None):
        self._mesh = mesh
        self._material = material_property(material_name, mesh.params, density)
        self._finite, self._symmetric = model_traits(mesh.model)
        names = _capi.lib().neon_model_variable_names(int(mesh.model)).decode()
        self._names = set(names.split(";")) if names else set()

    def update_internal_variables(self, time_step_size: float) -> None:
        """Update the internal variables and tangents at every quadrature point for the
        displacement the mesh holds (``neon_mesh_update_internal_variables``)."""
        self._mesh.update_internal_variables(time_step_size)

    def intrinsic_material(self) -> material_property:
        return self._material

    def is_finite_deformation(self) -> bool:
        return self._finite

    def is_symmetric(self) -> bool:
        return self._symmetric

    def variable_names(self) -> set:
        return self._names


class DeviceMesh:
    """Device-resident hex8 mesh, constitutive state and tangent (one per context)."""

    def __init__(self, X: np.ndarray, conn: np.ndarray, params: np.ndarray, model: int = MODEL_NEOHOOKE,
                 tangent: int = TANGENT_REFERENCE, context: Context | None = None, ageing=None, damage=None):
        """`ageing` = (scission, recombination) probabilities of the ageing microsphere model;
        `damage` = the six constants of the J2 damage model (``neon_mesh_set_damage``)."""
        self.ctx = context or Context(0)
        self.X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1)
        self.params = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        self.n_nodes = len(self.X) // 3
        self.n_elem = len(self.conn) // 8
        self.model = model
        self.ctx.check(_capi.lib().neon_mesh_create(
            self.ctx.handle, self.n_nodes, _capi.f64(self.X), self.n_elem, _capi.i32(self.conn), model,
            _capi.f64(self.params), tangent))
        if ageing is not None:
            self.set_ageing(*ageing)
        if damage is not None:
            self.set_damage(damage)
        self._constitutive = None

    def set_damage(self, constants) -> None:
        """The J2 damage model's material constants: kinematic hardening modulus, softening
        multiplier, plasticity viscous denominator and exponent, damage viscous denominator and
        exponent (``neon_mesh_set_damage``)."""
        c = np.ascontiguousarray(constants, dtype=np.float64)
        if c.size != 6:
            raise ValueError("six damage constants")
        self.ctx.check(_capi.lib().neon_mesh_set_damage(self.ctx.handle, _capi.f64(c)))

    def damage_state(self) -> dict:
        """back stress (Q, 3, 3), damage (Q), energy release rate (Q) of the J2 damage model."""
        Q = 8 * self.n_elem
        b, d, y = np.zeros(9 * Q), np.zeros(Q), np.zeros(Q)
        self.ctx.check(_capi.lib().neon_mesh_damage_state(self.ctx.handle, _capi.f64(b), _capi.f64(d), _capi.f64(y)))
        return dict(back_stress=b.reshape(Q, 3, 3), damage=d, energy=y)

    def constitutive(self, material_name: str = "", density: float | None = None) -> constitutive_model:
        """The mesh's constitutive model (``submesh::constitutive``, submesh.hpp:62)."""
        if self._constitutive is None or material_name or density is not None:
            self._constitutive = constitutive_model(self, material_name, density)
        return self._constitutive

    def is_symmetric(self) -> bool:
        """Whether the tangent is symmetric: every submesh's model is (mesh.cpp:47-53; one model
        per device mesh here)."""
        return model_traits(self.model)[1]

    def is_finite_deformation(self) -> bool:
        return model_traits(self.model)[0]

    def set_ageing(self, scission: float, recombination: float) -> None:
        """The ageing microsphere's scission and recombination probabilities
        (``neon_mesh_set_ageing``)."""
        self.ctx.check(_capi.lib().neon_mesh_set_ageing(self.ctx.handle, float(scission), float(recombination)))

    def ageing_state(self) -> np.ndarray:
        """(E*8, 11) network state of the ageing microsphere: active / inactive shear modulus,
        reduction factor, active / inactive segments, secondary intermediate stress (11, 22, 33,
        23, 13, 12)."""
        out = np.zeros((8 * self.n_elem, 11))
        self.ctx.check(_capi.lib().neon_mesh_ageing_state(self.ctx.handle, _capi.f64(out)))
        return out

    @property
    def n_dofs(self) -> int:
        return 3 * self.n_nodes

    def set_parameters(self, params: np.ndarray) -> None:
        self.params = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        self.ctx.check(_capi.lib().neon_mesh_set_parameters(self.ctx.handle, _capi.f64(self.params)))

    def pattern(self):
        nnz = C.c_int64(0)
        self.ctx.check(_capi.lib().neon_mesh_pattern(self.ctx.handle, C.byref(nnz), None, None))
        rp = np.zeros(self.n_dofs + 1, np.int32)
        ci = np.zeros(nnz.value, np.int32)
        self.ctx.check(_capi.lib().neon_mesh_pattern(self.ctx.handle, C.byref(nnz), _capi.i32(rp), _capi.i32(ci)))
        return rp, ci

    def nnz(self) -> int:
        nnz = C.c_int64(0)
        self.ctx.check(_capi.lib().neon_mesh_pattern(self.ctx.handle, C.byref(nnz), None, None))
        return nnz.value

    def set_displacement(self, u: np.ndarray) -> None:
        u = np.ascontiguousarray(u, dtype=np.float64)
        self.ctx.check(_capi.lib().neon_mesh_set_displacement(self.ctx.handle, _capi.f64(u)))

    def update_internal_variables(self, dt: float = 0.0) -> None:
        self.ctx.check(_capi.lib().neon_mesh_update_internal_variables(self.ctx.handle, dt))

    def save_internal_variables(self, converged: bool) -> None:
        self.ctx.check(_capi.lib().neon_mesh_save_internal_variables(self.ctx.handle, int(converged)))

    def sliced_info(self) -> dict:
        """Shape of the sliced tangent copy (``neon_mesh_sliced_info``)."""
        sl, idx = C.c_int32(0), C.c_int32(0)
        st, sh = C.c_int64(0), C.c_int64(0)
        self.ctx.check(_capi.lib().neon_mesh_sliced_info(self.ctx.handle, C.byref(sl), C.byref(st), C.byref(sh),
                                                         C.byref(idx)))
        return {"slices": sl.value, "steps": st.value, "shared_steps": sh.value, "index16": bool(idx.value)}

    def symmetric_info(self) -> dict:
        """Shape of the symmetric sliced copy (``neon_mesh_symmetric_info``): whether the solvers
        can use it, the pattern's distinct upper block offsets and the schedule's wave slots."""
        us, off, slots = C.c_int32(0), C.c_int32(0), C.c_int32(0)
        self.ctx.check(_capi.lib().neon_mesh_symmetric_info(self.ctx.handle, C.byref(us), C.byref(off),
                                                            C.byref(slots)))
        return {"usable": bool(us.value), "offsets": off.value, "slots": slots.value}

    def displacement(self) -> np.ndarray:
        """The displacement held by the mesh (``neon_mesh_displacement``)."""
        u = np.zeros(self.n_dofs)
        self.ctx.check(_capi.lib().neon_mesh_displacement(self.ctx.handle, _capi.f64(u)))
        return u

    def qp_state(self):
        Q = 8 * self.n_elem
        sigma, D = np.zeros(9 * Q), np.zeros(36 * Q)
        detF, eps_p, ebar, vm = np.zeros(Q), np.zeros(9 * Q), np.zeros(Q), np.zeros(Q)
        self.ctx.check(_capi.lib().neon_mesh_qp_state(self.ctx.handle, _capi.f64(sigma), _capi.f64(D), _capi.f64(detF),
                                                      _capi.f64(eps_p), _capi.f64(ebar), _capi.f64(vm)))
        return dict(sigma=sigma.reshape(Q, 3, 3), D=D.reshape(Q, 6, 6), detF=detF, eps_p=eps_p.reshape(Q, 3, 3),
                    ebar=ebar, vm=vm)

    def assemble(self, precision: int = PRECISION_FP64) -> None:
        self.ctx.check(_capi.lib().neon_mesh_assemble(self.ctx.handle, precision))

    def mass(self, density) -> np.ndarray:
        """The consistent mass (``submesh::consistent_mass``) in the pattern's CSR order; density
        per element or one value."""
        rho = np.ascontiguousarray(np.broadcast_to(np.asarray(density, dtype=np.float64), (self.n_elem,)))
        val = np.zeros(self.nnz())
        self.ctx.check(_capi.lib().neon_mesh_mass(self.ctx.handle, _capi.f64(rho), _capi.f64(val)))
        return val

    def download(self, stiffness: bool = True, internal_force: bool = True):
        val = np.zeros(self.nnz()) if stiffness else None
        f = np.zeros(self.n_dofs) if internal_force else None
        self.ctx.check(_capi.lib().neon_mesh_download(self.ctx.handle, _capi.f64(val), _capi.f64(f)))
        return val, f

    def enforce_dirichlet(self, fixed: np.ndarray, b: np.ndarray | None = None) -> None:
        fixed = np.ascontiguousarray(fixed, dtype=np.int32)
        self.ctx.check(_capi.lib().neon_mesh_enforce_dirichlet(self.ctx.handle, len(fixed), _capi.i32(fixed),
                                                               _capi.f64(b)))

    def spmv_config(self, lanes: int = 32,
                    flags: int = _capi.SPMV_SLICED | _capi.SPMV_NONTEMPORAL | _capi.SPMV_SYMMETRIC,
                    blocks_fp64: int = 0, blocks_fp32: int = 0) -> None:
        """Shape of the node-block SpMV used by the solvers on this mesh (tuning knob; the
        defaults are the library's: the sliced copy, one wavefront per 64 nodes, values read with
        nontemporal loads, the symmetric copy where the tangent allows it)."""
        self.ctx.check(_capi.lib().neon_mesh_spmv_config(self.ctx.handle, lanes, flags, blocks_fp64, blocks_fp32))

    def adjoint_gradient(self, dJdu: np.ndarray, fixed: np.ndarray, tol: float = 1e-12,
                         max_iterations: int = 20000):
        """dJ/dparams (n_elem x 4) at the held state for an objective with gradient dJdu,
        through the adjoint solve K^T lambda = dJ/du (consistent tangent, neo-Hooke)."""
        dJdu = np.ascontiguousarray(dJdu, dtype=np.float64)
        fixed = np.ascontiguousarray(fixed, dtype=np.int32)
        grad = np.zeros((self.n_elem, 4))
        its = C.c_int32(0)
        code = _capi.lib().neon_adjoint_gradient(self.ctx.handle, _capi.f64(dJdu), len(fixed), _capi.i32(fixed), tol,
                                                 max_iterations, _capi.f64(grad), C.byref(its))
        self.adjoint_iterations = its.value
        self.ctx.check(code)
        return grad

    def solve(self, b: np.ndarray, solver: int = SOLVER_CG, criterion: int = STOP_RELATIVE, tol: float = 1e-10,
              max_iterations: int = 10000):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n_dofs)
        its, err = C.c_int32(0), C.c_double(0.0)
        code = _capi.lib().neon_mesh_solve(self.ctx.handle, solver, criterion, tol, max_iterations, _capi.f64(b),
                                           _capi.f64(x), C.byref(its), C.byref(err))
        self.ctx.check(code)
        return x, its.value, err.value


def _bc_structs(boundaries):
    """boundaries: list of dicts {dofs, times, values} -> ctypes array (+ keepalive)."""
    keep = []
    arr = (_capi.Dirichlet * max(1, len(boundaries)))()
    for k, bc in enumerate(boundaries):
        d = np.ascontiguousarray(bc["dofs"], dtype=np.int32)
        t = np.ascontiguousarray(bc["times"], dtype=np.float64)
        v = np.ascontiguousarray(bc["values"], dtype=np.float64)
        keep += [d, t, v]
        arr[k] = _capi.Dirichlet(len(d), _capi.i32(d), len(t), _capi.f64(t), _capi.f64(v))
    return arr, keep


def _load_structs(loads):
    """loads: list of dicts {dofs, values, times, loads[, fixed_time]} -> ctypes array (+ keepalive)."""
    keep = []
    arr = (_capi.Neumann * max(1, len(loads)))()
    for k, ld in enumerate(loads):
        d = np.ascontiguousarray(ld["dofs"], dtype=np.int32)
        v = np.ascontiguousarray(ld["values"], dtype=np.float64)
        t = np.ascontiguousarray(ld["times"], dtype=np.float64)
        q = np.ascontiguousarray(ld["loads"], dtype=np.float64)
        if len(d) != len(v) or len(t) != len(q):
            raise ValueError("load dofs/values or times/loads differ in length")
        keep += [d, v, t, q]
        arr[k] = _capi.Neumann(len(d), _capi.i32(d), _capi.f64(v), len(t), _capi.f64(t), _capi.f64(q),
                               float(ld.get("fixed_time", -1.0)))
    return arr, keep


def surface_load_vector(X: np.ndarray, faces: np.ndarray, component: int, reduced: bool = False) -> np.ndarray:
    """Unit traction (component 0..2) or unit pressure (component -1) load vector (3N) of
    quadrilateral faces on the initial configuration (``neon_surface_load_vector``)."""
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 3)
    faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 4)
    out = np.zeros(3 * len(X))
    code = _capi.lib().neon_surface_load_vector(len(X), _capi.f64(X), len(faces), _capi.i32(faces), int(component),
                                                int(bool(reduced)), _capi.f64(out))
    _raise(code, None)
    return out


def volume_load_vector(X: np.ndarray, conn: np.ndarray, component: int, reduced: bool = False) -> np.ndarray:
    """Unit body-force load vector (3N) of hex8 elements (``neon_volume_load_vector``)."""
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 3)
    conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1, 8)
    out = np.zeros(3 * len(X))
    code = _capi.lib().neon_volume_load_vector(len(X), _capi.f64(X), len(conn), _capi.i32(conn), int(component),
                                               int(bool(reduced)), _capi.f64(out))
    _raise(code, None)
    return out


class static_matrix:
    """Newton driver (``static_matrix::solve``, src/assembler/mechanics/static_matrix.hpp:145-467)
    on a :class:`DeviceMesh`.  ``simulation`` uses neon's keys: ``time``
    {period, increments{initial, minimum, maximum, adaptive}}, ``nonlinear_options``
    {displacement_tolerance, residual_tolerance, linear_iterations, absolute_tolerance},
    ``linear_solver`` {tolerance, maximum_iterations, backend{precision}}; plus ``criterion``,
    ``norm_mode``, ``max_load_steps``, ``max_total_iterations`` and ``record_trajectory`` (keep the
    load history on the device for the J2 adjoint) extensions.  ``loads``: nonfollower loads
    (``neon_neumann``), each {dofs, values (unit-load vector), times, loads[, fixed_time]}."""

    def __init__(self, mesh: DeviceMesh, simulation: dict, boundaries: list, loads: list | None = None):
        self.mesh = mesh
        self.loads = list(loads or [])
        nl = simulation["nonlinear_options"]
        if "displacement_tolerance" not in nl:
            raise ValueError("displacement_tolerance not specified in nonlinear_options")
        if "residual_tolerance" not in nl:
            raise ValueError("residual_tolerance not specified in nonlinear_options")
        t = simulation["time"]
        inc = t["increments"]
        ls = simulation.get("linear_solver", {})
        backend = ls.get("backend", {})
        o = _capi.NewtonOptions()
        o.period = t["period"]
        o.initial = inc["initial"]
        o.minimum = inc["minimum"]
        o.maximum = inc["maximum"]
        o.adaptive = int(bool(inc["adaptive"]))
        o.residual_tolerance = nl["residual_tolerance"]
        o.displacement_tolerance = nl["displacement_tolerance"]
        o.max_newton_iterations = nl.get("linear_iterations", 10)
        o.absolute_norm = int("absolute_tolerance" in nl)
        o.norm_mode = simulation.get("norm_mode", 0)
        o.solver_tolerance = ls.get("tolerance", 1.0e-5)
        o.solver_max_iterations = ls.get("maximum_iterations", 2000)
        o.solver_criterion = simulation.get("criterion", STOP_RELATIVE)
        # make_linear_solver(linear_solver, mesh.is_symmetric()) (static_matrix.hpp:109-112,
        # linear_solver_factory.cpp:150-161): conjugate gradient for a symmetric tangent,
        # BiCGSTAB otherwise
        o.solver = SOLVER_CG if mesh.is_symmetric() else SOLVER_BICGSTAB
        o.precision = {"fp64": PRECISION_FP64, "mixed": PRECISION_MIXED}[backend.get("precision", "fp64")]
        o.max_load_steps = simulation.get("max_load_steps", 0)
        o.max_total_iterations = simulation.get("max_total_iterations", 0)
        o.record_trajectory = int(bool(simulation.get("record_trajectory", False)))
        self.options = o
        self.boundaries = boundaries
        self.records: list[dict] = []
        self.timing: dict = {}
        self.latin_search_direction = None

    def solve(self, capacity: int = 4096, on_step=None) -> np.ndarray:
        """Runs every load step.  ``on_step(step, time)`` is called where neon writes its
        output (the initial state, after each cut-back and after each converged step); it may
        read the mesh (``DeviceMesh.displacement``, ``qp_state``, ``download``) but not change it."""
        m = self.mesh
        u = np.zeros(m.n_dofs)
        failure = []

        def _cb(user, step, time):
            # exceptions cannot cross the C callback: keep the first and re-raise after the solve
            if failure:
                return
            try:
                on_step(int(step), float(time))
            except BaseException as e:  # noqa: BLE001
                failure.append(e)

        cb = _capi.StepCallback(_cb) if on_step is not None else _capi.StepCallback()
        self.options.on_step = cb
        recs = (_capi.NewtonRecord * capacity)()
        nrec = C.c_int32(0)
        tm = _capi.NewtonTiming()
        bcs, keep = _bc_structs(self.boundaries)
        if self.latin_search_direction is not None:
            lds, keep_l = _load_structs(self.loads) if self.loads else (None, None)
            code = _capi.lib().neon_latin_solve(m.ctx.handle, C.byref(self.options), self.latin_search_direction,
                                                len(self.boundaries), bcs, len(self.loads), lds, _capi.f64(u), recs,
                                                capacity, C.byref(nrec), C.byref(tm))
            del keep_l
        elif self.loads:
            lds, keep_l = _load_structs(self.loads)
            code = _capi.lib().neon_newton_solve_loaded(m.ctx.handle, C.byref(self.options), len(self.boundaries),
                                                        bcs, len(self.loads), lds, _capi.f64(u), recs, capacity,
                                                        C.byref(nrec), C.byref(tm))
            del keep_l
        else:
            code = _capi.lib().neon_newton_solve(m.ctx.handle, C.byref(self.options), len(self.boundaries), bcs,
                                                 _capi.f64(u), recs, capacity, C.byref(nrec), C.byref(tm))
        del keep
        self.records = [
            dict(step=r.step, iteration=r.iteration, solver_iterations=r.solver_iterations, status=r.status,
                 step_time=r.step_time, displacement_norm=r.displacement_norm, force_norm=r.force_norm)
            for r in recs[: min(nrec.value, capacity)]
        ]
        self.timing = {name: getattr(tm, name) for name, _ in _capi.NewtonTiming._fields_}
        self.options.on_step = _capi.StepCallback()
        del cb
        if failure:
            raise failure[0]
        m.ctx.check(code)
        return u


class latin_matrix(static_matrix):
    """LATIN driver (``latin_matrix``, src/assembler/mechanics/latin_matrix.hpp:150-244): the
    load stepping and norms of :class:`static_matrix`; after each step's first iteration the
    right-hand side is the LATIN force of ``latin_submesh::incremental_latin_internal_force``
    (latin_submesh.hpp:62-111) with ``nonlinear_options["latin_search_direction"]``."""

    def __init__(self, mesh: DeviceMesh, simulation: dict, boundaries: list, loads: list | None = None):
        super().__init__(mesh, simulation, boundaries, loads)
        nl = simulation["nonlinear_options"]
        if "latin_search_direction" not in nl:
            raise ValueError("latin_search_direction not specified in nonlinear_options")
        self.latin_search_direction = float(nl["latin_search_direction"])


def make_eigen_solver(data: dict) -> dict:
    """``make_eigen_solver`` (src/solver/eigen/eigen_solver_factory.cpp:12-77): the number of
    eigenvalues (``"eigenvalues"``, default 10) and the end of the spectrum.  ``"arpack"`` maps to
    the device solver; neon's ``"lanczos"`` and ``"power_iteration"`` need its OpenCL build and
    solve no generalised problem (power_iteration.cpp:45-48, lanczos_ocl.cpp:49), so they are
    rejected as neon's default build rejects them.  Without ``"spectrum"`` neon leaves the choice
    uninitialised; the lower end is taken here."""
    if "type" not in data:
        raise ValueError('Eigen solver type was not provided.  Please use "power_iteration", "arpack" or "lanczos"')
    values = 10
    if "eigenvalues" in data:
        values = int(data["eigenvalues"])
        if values <= 0:
            raise ValueError("The number of requested eigenvalues must be positive")
    spectrum = SPECTRUM_LOWER
    if "spectrum" in data:
        if data["spectrum"] == "lower":
            spectrum = SPECTRUM_LOWER
        elif data["spectrum"] == "upper":
            spectrum = SPECTRUM_UPPER
        else:
            raise ValueError('A spectrum was requested but only "lower" and "upper" options are supported')
    kind = data["type"]
    if kind in ("power_iteration", "lanczos"):
        raise ValueError(f'The "{kind}" option requires OpenCL support.  Recompile with -DENABLE_OPENCL=1')
    if kind != "arpack":
        raise ValueError(f'Eigen solver type "{kind}" is not recognised.  Please use "arpack"')
    return {"type": kind, "values": values, "spectrum": spectrum}


class natural_frequency_matrix:
    """``natural_frequency_matrix`` (src/assembler/mechanics/natural_frequency_matrix.hpp:17-161) on
    a :class:`DeviceMesh`: K at the mesh's held state and the consistent mass, neon's homogeneous
    Dirichlet treatment of both (the dofs of ``boundaries``), and the eigenvalues of
    K x = lambda M x on the device (``neon_natural_frequency``).  ``solve`` returns the eigenvalues
    (increasing) and the M-orthonormal mode shapes (3N x values); ``frequencies`` are their square
    roots (rad / time), as neon writes them."""

    def __init__(self, mesh: DeviceMesh, eigen_solver: dict, boundaries: list, density):
        self.mesh = mesh
        self.solver = make_eigen_solver(eigen_solver)
        d = [np.asarray(b["dofs"]) for b in boundaries]
        self.fixed = np.ascontiguousarray(np.unique(np.concatenate(d)) if d else np.zeros(0), dtype=np.int32)
        self.density = np.ascontiguousarray(np.broadcast_to(np.asarray(density, dtype=np.float64), (mesh.n_elem,)))
        self.eigenvalues = None
        self.eigenvectors = None
        self.solves = 0

    def solve(self, tolerance: float = 1.0e-8, max_basis: int = 0, block: int = 0):
        """tolerance: on every wanted Ritz pair's residual |OP y - theta y|_M relative to theta (the
        eigenvalues' error goes with its square)."""
        k = self.solver["values"]
        vals = np.zeros(k)
        vecs = np.zeros((k, self.mesh.n_dofs))
        solves = C.c_int32(0)
        self.mesh.ctx.check(_capi.lib().neon_natural_frequency(
            self.mesh.ctx.handle, _capi.f64(self.density), len(self.fixed), _capi.i32(self.fixed), k,
            self.solver["spectrum"], tolerance, max_basis, block, _capi.f64(vals), _capi.f64(vecs), C.byref(solves)))
        self.eigenvalues, self.eigenvectors, self.solves = vals, vecs.T.copy(), solves.value
        return self.eigenvalues, self.eigenvectors

    @property
    def frequencies(self) -> np.ndarray:
        return np.sqrt(self.eigenvalues)

    def eigenvalue_table(self) -> str:
        """The table of natural_frequency_matrix::print_eigenvalue_table (:139-160)."""
        lines = ["      Eigenvalue | radians / time | cycles / time", "      " + "-" * 43]
        for i, lam in enumerate(self.eigenvalues):
            w = np.sqrt(lam)
            lines.append(f"      {i + 1:>10}{w:>17.6g}{w / (2.0 * np.pi):>16.6g}")
        return "\n".join(lines)
