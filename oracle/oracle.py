"""ctypes binding of the CPU FP64 oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU baseline of the
benchmark.  Only tests, the smoke check and the benchmark's CPU-baseline leg import it.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

P = C.POINTER
i32p, f64p = P(C.c_int32), P(C.c_double)

OK, NONCONVERGED, NAN_OR_INF, NEGATIVE_JACOBIAN, RETURN_MAP_FAILED, BREAKDOWN = 0, 1, 2, 3, 4, 5
FATAL, BAD_ARGUMENT = -1, -2
MODEL_NEOHOOKE, MODEL_J2, MODEL_LINEAR_ELASTIC = 0, 1, 2
MODEL_MICROSPHERE_GAUSSIAN, MODEL_MICROSPHERE_LANGEVIN, MODEL_MICROSPHERE_NONAFFINE = 3, 4, 5
MODEL_MICROSPHERE_AGEING = 6
MODEL_J2_DAMAGE = 7
SPHERE_BO33 = 0x100
STOP_RELATIVE, STOP_ABSOLUTE_PRECONDITIONED = 0, 1
TANGENT_REFERENCE, TANGENT_CONSISTENT = 0, 1


class TimeOptions(C.Structure):
    _fields_ = [("period", C.c_double), ("initial", C.c_double), ("minimum", C.c_double), ("maximum", C.c_double),
                ("adaptive", C.c_int32)]


class Dirichlet(C.Structure):
    _fields_ = [("n_dofs", C.c_int32), ("dofs", i32p), ("n_points", C.c_int32), ("times", f64p),
                ("values", f64p)]


class Neumann(C.Structure):
    _fields_ = [("n_entries", C.c_int32), ("dofs", i32p), ("values", f64p), ("n_points", C.c_int32),
                ("times", f64p), ("loads", f64p), ("fixed_time", C.c_double)]


class NewtonOptions(C.Structure):
    _fields_ = [("time", TimeOptions), ("residual_tolerance", C.c_double), ("displacement_tolerance", C.c_double),
                ("max_newton_iterations", C.c_int32), ("absolute_norm", C.c_int32), ("norm_mode", C.c_int32),
                ("solver_tolerance", C.c_double), ("solver_max_iterations", C.c_int32),
                ("solver_criterion", C.c_int32), ("solver", C.c_int32), ("threads", C.c_int32),
                ("rcm", C.c_int32), ("record_trajectory", C.c_int32), ("latin", C.c_int32),
                ("latin_search_direction", C.c_double)]


class NewtonRecord(C.Structure):
    _fields_ = [("step", C.c_int32), ("iteration", C.c_int32), ("solver_iterations", C.c_int32),
                ("status", C.c_int32), ("step_time", C.c_double), ("displacement_norm", C.c_double),
                ("force_norm", C.c_double)]


_SIG = {
    "orc_hex8_tables": [f64p, f64p, f64p],
    "orc_neohooke_qp": [f64p, C.c_double, C.c_double, f64p, f64p],
    "orc_microsphere_qp": [C.c_int32, f64p, f64p, C.c_int32, f64p, f64p],
    "orc_sphere_rule": [C.c_int32, f64p, f64p],
    "orc_model_set_ageing": [C.c_void_p, C.c_double, C.c_double],
    "orc_get_ageing_state": [C.c_void_p, f64p],
    "orc_model_set_damage": [C.c_void_p, f64p],
    "orc_j2_damage_system": [f64p, f64p, C.c_double, C.c_double, C.c_double, C.c_double, f64p, C.c_double, f64p,
                             f64p, f64p],
    "orc_get_damage_state": [C.c_void_p, f64p, f64p, f64p],
    "orc_j2_damage_qp": [f64p, C.c_double, C.c_double, C.c_double, f64p, C.c_double, f64p, f64p, f64p, f64p, f64p,
                         f64p, f64p, f64p, f64p, i32p],
    "orc_j2_qp": [f64p, C.c_double, C.c_double, C.c_double, C.c_double, f64p, f64p, f64p, f64p, f64p],
    "orc_model_create": [C.c_int32, f64p, C.c_int32, i32p, C.c_int32, f64p, C.c_int32, P(C.c_void_p)],
    "orc_model_destroy": [C.c_void_p],
    "orc_update_internal_variables": [C.c_void_p, f64p, C.c_double],
    "orc_save_internal_variables": [C.c_void_p, C.c_int32],
    "orc_get_qp_state": [C.c_void_p, f64p, f64p, f64p, f64p, f64p, f64p],
    "orc_element_matrices": [C.c_void_p, C.c_int32, f64p, f64p],
    "orc_parameter_gradient": [C.c_void_p, f64p, f64p],
    "orc_adjoint_j2": [C.c_void_p, f64p, C.c_int32, i32p, C.c_double, f64p],
    "orc_trajectory_steps": [C.c_void_p],
    "orc_sparsity_fill": [C.c_void_p, i32p, i32p],
    "orc_assemble": [C.c_void_p, i32p, i32p, f64p, f64p, C.c_int32],
    "orc_element_mass": [C.c_void_p, C.c_int32, C.c_double, f64p],
    "orc_assemble_mass": [C.c_void_p, i32p, i32p, f64p, f64p],
    "orc_enforce_dirichlet": [C.c_int32, i32p, i32p, f64p, f64p, C.c_int32, i32p],
    "orc_latin_force": [C.c_void_p, C.c_double, f64p],
    "orc_spmv": [C.c_int32, i32p, i32p, f64p, f64p, f64p, C.c_int32],
    "orc_rcm": [C.c_int32, i32p, i32p, i32p],
    "orc_time_create": [P(TimeOptions), C.c_int32, f64p, P(C.c_void_p)],
    "orc_time_destroy": [C.c_void_p],
    "orc_time_update": [C.c_void_p, C.c_int32],
    "orc_time_query": [C.c_void_p, f64p, f64p, i32p, i32p],
    "orc_newton_solve": [C.c_void_p, P(NewtonOptions), C.c_int32, P(Dirichlet), f64p, P(NewtonRecord), C.c_int32,
                         i32p],
    "orc_newton_solve_loaded": [C.c_void_p, P(NewtonOptions), C.c_int32, P(Dirichlet), C.c_int32, P(Neumann), f64p,
                                P(NewtonRecord), C.c_int32, i32p],
}
for _k in ("orc_cg_eigen", "orc_cg_cuda", "orc_bicgstab_eigen", "orc_bicgstab_cuda"):
    _SIG[_k] = [C.c_int32, i32p, i32p, f64p, f64p, f64p, C.c_double, C.c_int32, C.c_int32, i32p, f64p]

_lib = None


def build() -> pathlib.Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        for name, args in _SIG.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = C.c_int
        h.orc_boundary_value.argtypes = [C.c_int32, f64p, f64p, C.c_double]
        h.orc_boundary_value.restype = C.c_double
        h.orc_sparsity_nnz.argtypes = [C.c_void_p]
        h.orc_sparsity_nnz.restype = C.c_int64
        _lib = h
    return _lib


def _f(a):
    return None if a is None else a.ctypes.data_as(f64p)


def _i(a):
    return None if a is None else a.ctypes.data_as(i32p)


def hex8_tables():
    N, dN, w = np.zeros((8, 8)), np.zeros((8, 8, 3)), np.zeros(8)
    lib().orc_hex8_tables(_f(N), _f(dN), _f(w))
    return N, dN, w


def neohooke_qp(F, bulk, shear):
    F = np.ascontiguousarray(F, dtype=np.float64)
    s, D = np.zeros((3, 3)), np.zeros((6, 6))
    lib().orc_neohooke_qp(_f(F), bulk, shear, _f(s), _f(D))
    return s, D


def microsphere_qp(model, F, params, consistent=False):
    """sigma (3x3) and D (6x6) of a microsphere model (id | rule flag) at one QP;
    params = (K, G, segments per chain, non-affine stretch parameter)."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    prm = np.ascontiguousarray(params, dtype=np.float64)
    s, D = np.zeros((3, 3)), np.zeros((6, 6))
    code = lib().orc_microsphere_qp(model, _f(F), _f(prm), int(consistent), _f(s), _f(D))
    if code != OK:
        raise RuntimeError(f"orc_microsphere_qp failed: {code}")
    return s, D


def sphere_rule(model):
    """(weights, directions) of the unit-sphere rule of a microsphere model id."""
    w, r = np.zeros(33), np.zeros((33, 3))
    n = lib().orc_sphere_rule(model, _f(w), _f(r))
    return w[:n], r[:n]


def j2_qp(H, bulk, shear, yield_stress, hardening, eps_p=None, ebar=0.0):
    H = np.ascontiguousarray(H, dtype=np.float64)
    ep = np.zeros((3, 3)) if eps_p is None else np.array(eps_p, dtype=np.float64)
    eb = np.array([ebar], dtype=np.float64)
    s, D, vm = np.zeros((3, 3)), np.zeros((6, 6)), np.zeros(1)
    code = lib().orc_j2_qp(_f(H), bulk, shear, yield_stress, hardening, _f(ep), _f(eb), _f(s), _f(D), _f(vm))
    return dict(status=code, sigma=s, D=D, vm=vm[0], eps_p=ep, ebar=eb[0])


def j2_damage_qp(H, bulk, shear, yield_stress, constants, dt, state=None):
    """One QP of small_strain_J2_plasticity_damage::update_internal_variables.  constants =
    (kinematic hardening modulus, softening multiplier, plasticity viscous denominator and
    exponent, damage viscous denominator and exponent); state: dict of eps_p, ebar, back_stress,
    damage, energy (updated in place and returned with sigma, D, vm, kinematic_hardening)."""
    H = np.ascontiguousarray(H, dtype=np.float64)
    st = dict(eps_p=np.zeros((3, 3)), ebar=np.zeros(1), back_stress=np.zeros((3, 3)), damage=np.zeros(1),
              energy=np.zeros(1)) if state is None else state
    for k in ("ebar", "damage", "energy"):
        st[k] = np.atleast_1d(np.asarray(st[k], dtype=np.float64)).copy()
    for k in ("eps_p", "back_stress"):
        st[k] = np.ascontiguousarray(st[k], dtype=np.float64).copy()
    c = np.ascontiguousarray(constants, dtype=np.float64)
    s, D, vm, X, its = np.zeros((3, 3)), np.zeros((6, 6)), np.zeros(1), np.zeros((3, 3)), C.c_int32(0)
    code = lib().orc_j2_damage_qp(_f(H), bulk, shear, yield_stress, _f(c), dt, _f(st["eps_p"]), _f(st["ebar"]),
                                  _f(st["back_stress"]), _f(st["damage"]), _f(st["energy"]), _f(X), _f(s), _f(D),
                                  _f(vm), C.byref(its))
    st.update(status=code, sigma=s, D=D, vm=vm[0], kinematic_hardening=X, iterations=its.value)
    return st


def j2_damage_system(eps_e_t, back_stress, damage, bulk, shear, yield_stress, constants, dt, y):
    """Residual (16) and Jacobian (16x16) of the damage model's return map at unknowns y."""
    r, J = np.zeros(16), np.zeros((16, 16))
    lib().orc_j2_damage_system(_f(np.ascontiguousarray(eps_e_t, dtype=np.float64)),
                               _f(np.ascontiguousarray(back_stress, dtype=np.float64)), float(damage), bulk, shear,
                               yield_stress, _f(np.ascontiguousarray(constants, dtype=np.float64)), dt,
                               _f(np.ascontiguousarray(y, dtype=np.float64)), _f(r), _f(J))
    return r, J


class Model:
    """Mesh-level oracle model (constitutive state + element matrices + assembly)."""

    def __init__(self, X, conn, params, model=MODEL_NEOHOOKE, tangent=TANGENT_REFERENCE):
        self.X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1)
        self.params = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        self.n_nodes, self.n_elem = len(self.X) // 3, len(self.conn) // 8
        self.h = C.c_void_p()
        code = lib().orc_model_create(self.n_nodes, _f(self.X), self.n_elem, _i(self.conn), model, _f(self.params),
                                      tangent, C.byref(self.h))
        if code != OK:
            raise RuntimeError(f"orc_model_create failed: {code}")

    def __del__(self):
        try:
            if self.h:
                lib().orc_model_destroy(self.h)
        except Exception:
            pass

    @property
    def n_dofs(self):
        return 3 * self.n_nodes

    def update(self, u, dt=0.0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return lib().orc_update_internal_variables(self.h, _f(u), dt)

    def save(self, converged):
        lib().orc_save_internal_variables(self.h, int(converged))

    def qp_state(self):
        Q = 8 * self.n_elem
        s, D, J, ep, eb, vm = (np.zeros(9 * Q), np.zeros(36 * Q), np.zeros(Q), np.zeros(9 * Q), np.zeros(Q),
                               np.zeros(Q))
        lib().orc_get_qp_state(self.h, _f(s), _f(D), _f(J), _f(ep), _f(eb), _f(vm))
        return dict(sigma=s.reshape(Q, 3, 3), D=D.reshape(Q, 6, 6), detF=J, eps_p=ep.reshape(Q, 3, 3), ebar=eb, vm=vm)

    def element(self, e):
        ke, fe = np.zeros((24, 24)), np.zeros(24)
        lib().orc_element_matrices(self.h, e, _f(ke), _f(fe))
        return ke, fe

    def parameter_gradient(self, lam):
        """-lambda_e . d f_e / d params_e (n_elem x 4) at the current state (neo-Hooke)."""
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        g = np.zeros((self.n_elem, 4))
        code = lib().orc_parameter_gradient(self.h, _f(lam), _f(g))
        if code != OK:
            raise RuntimeError(f"orc_parameter_gradient failed: {code}")
        return g

    def adjoint_j2(self, dJdu, fixed, tolerance=1e-14):
        """dJ/dparams (n_elem x 4) over the load history recorded by the last newton() with
        record_trajectory (J2, consistent mode)."""
        dJdu = np.ascontiguousarray(dJdu, dtype=np.float64)
        fixed = np.ascontiguousarray(fixed, dtype=np.int32)
        g = np.zeros((self.n_elem, 4))
        code = lib().orc_adjoint_j2(self.h, _f(dJdu), len(fixed), _i(fixed), tolerance, _f(g))
        if code != OK:
            raise RuntimeError(f"orc_adjoint_j2 failed: {code}")
        return g

    def trajectory_steps(self):
        return lib().orc_trajectory_steps(self.h)

    def pattern(self):
        nnz = lib().orc_sparsity_nnz(self.h)
        rp, ci = np.zeros(self.n_dofs + 1, np.int32), np.zeros(nnz, np.int32)
        lib().orc_sparsity_fill(self.h, _i(rp), _i(ci))
        return rp, ci

    def assemble(self, rp, ci, threads=0):
        val, f = np.zeros(len(ci)), np.zeros(self.n_dofs)
        lib().orc_assemble(self.h, _i(rp), _i(ci), _f(val), _f(f), threads)
        return val, f

    def set_ageing(self, scission, recombination):
        if lib().orc_model_set_ageing(self.h, scission, recombination) != OK:
            raise RuntimeError("orc_model_set_ageing failed")

    def ageing_state(self):
        """(E*8, 11): active / inactive shear modulus, reduction factor, active / inactive
        segments, secondary intermediate stress (11, 22, 33, 23, 13, 12)."""
        out = np.zeros((8 * self.n_elem, 11))
        if lib().orc_get_ageing_state(self.h, _f(out)) != OK:
            raise RuntimeError("orc_get_ageing_state failed")
        return out

    def set_damage(self, constants):
        c = np.ascontiguousarray(constants, dtype=np.float64)
        if lib().orc_model_set_damage(self.h, _f(c)) != OK:
            raise RuntimeError("orc_model_set_damage failed")

    def damage_state(self):
        """back stress (Q, 3, 3), damage (Q), energy release rate (Q) of the J2 damage model."""
        Q = 8 * self.n_elem
        b, d, y = np.zeros(9 * Q), np.zeros(Q), np.zeros(Q)
        if lib().orc_get_damage_state(self.h, _f(b), _f(d), _f(y)) != OK:
            raise RuntimeError("orc_get_damage_state failed")
        return dict(back_stress=b.reshape(Q, 3, 3), damage=d, energy=y)

    def latin_force(self, search_direction):
        """orc_latin_force at the model's state (3N)."""
        f = np.zeros(3 * self.n_nodes)
        if lib().orc_latin_force(self.h, float(search_direction), _f(f)) != OK:
            raise RuntimeError("orc_latin_force failed")
        return f

    def element_mass(self, e, density):
        me = np.zeros((24, 24))
        lib().orc_element_mass(self.h, e, density, _f(me))
        return me

    def assemble_mass(self, rp, ci, density):
        """Consistent mass in the CSR pattern; density per element (or one value)."""
        rho = np.ascontiguousarray(np.broadcast_to(np.asarray(density, dtype=np.float64), (self.n_elem,)))
        val = np.zeros(len(ci))
        lib().orc_assemble_mass(self.h, _i(rp), _i(ci), _f(rho), _f(val))
        return val

    def newton(self, options: NewtonOptions, boundaries, capacity=4096, loads=None):
        """Newton driver; ``loads``: nonfollower loads {dofs, values, times, loads[, fixed_time]}."""
        arr = (Dirichlet * max(1, len(boundaries)))()
        keep = []
        for k, bc in enumerate(boundaries):
            d = np.ascontiguousarray(bc["dofs"], dtype=np.int32)
            t = np.ascontiguousarray(bc["times"], dtype=np.float64)
            v = np.ascontiguousarray(bc["values"], dtype=np.float64)
            keep += [d, t, v]
            arr[k] = Dirichlet(len(d), _i(d), len(t), _f(t), _f(v))
        loads = list(loads or [])
        larr = (Neumann * max(1, len(loads)))()
        for k, ld in enumerate(loads):
            d = np.ascontiguousarray(ld["dofs"], dtype=np.int32)
            v = np.ascontiguousarray(ld["values"], dtype=np.float64)
            t = np.ascontiguousarray(ld["times"], dtype=np.float64)
            q = np.ascontiguousarray(ld["loads"], dtype=np.float64)
            keep += [d, v, t, q]
            larr[k] = Neumann(len(d), _i(d), _f(v), len(t), _f(t), _f(q), float(ld.get("fixed_time", -1.0)))
        u = np.zeros(self.n_dofs)
        recs = (NewtonRecord * capacity)()
        n = C.c_int32(0)
        code = lib().orc_newton_solve_loaded(self.h, C.byref(options), len(boundaries), arr, len(loads), larr, _f(u),
                                             recs, capacity, C.byref(n))
        records = [dict(step=r.step, iteration=r.iteration, solver_iterations=r.solver_iterations, status=r.status,
                        step_time=r.step_time, displacement_norm=r.displacement_norm, force_norm=r.force_norm)
                   for r in recs[: min(n.value, capacity)]]
        return code, u, records


# ------------------------------------------------------------------ nonfollower load vectors
_GAUSS = 1.0 / np.sqrt(3.0)
_QUAD_NODES = np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]])  # quadrilateral.cpp:21-22
_HEX_NODES = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                       [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)  # hexahedron.cpp:21-28


def surface_load_vector(X, faces, component, reduced=False):
    """Unit-load vector (3N) of quadrilateral4 faces on the initial configuration X: traction in
    `component` (neumann.hpp:104-118: sum_l w N_a |x_xi x x_eta|) or, for component -1, a unit
    pressure (pressure.cpp:8-31: -sum_l w N_a n |x_xi x x_eta|).  2x2 Gauss points of weight 1,
    or the centre with weight 4 when reduced (quadrilateral_quadrature.cpp:10-24)."""
    X = np.asarray(X, dtype=np.float64).reshape(-1, 3)
    faces = np.asarray(faces, dtype=np.int64).reshape(-1, 4)
    pts = [(0.0, 0.0, 4.0)] if reduced else [(s * _GAUSS, t * _GAUSS, 1.0) for s, t in _QUAD_NODES]
    out = np.zeros(3 * len(X))
    xf = X[faces]  # (F, 4, 3)
    for xi, eta, w in pts:
        N = 0.25 * (1 + _QUAD_NODES[:, 0] * xi) * (1 + _QUAD_NODES[:, 1] * eta)
        dxi = 0.25 * _QUAD_NODES[:, 0] * (1 + _QUAD_NODES[:, 1] * eta)
        deta = 0.25 * _QUAD_NODES[:, 1] * (1 + _QUAD_NODES[:, 0] * xi)
        t1 = np.einsum("fai,a->fi", xf, dxi)
        t2 = np.einsum("fai,a->fi", xf, deta)
        c = np.cross(t1, t2)
        j = np.linalg.norm(c, axis=1)
        for a in range(4):
            if component >= 0:
                np.add.at(out, 3 * faces[:, a] + component, w * N[a] * j)
            else:
                for i in range(3):
                    np.add.at(out, 3 * faces[:, a] + i, -w * N[a] * c[:, i])
    return out


def volume_load_vector(X, conn, component, reduced=False):
    """Unit body-force vector (3N) of hex8 elements (neumann.hpp:159-173: sum_l w N_a det J),
    2x2x2 Gauss points or the centre with weight 8 when reduced."""
    X = np.asarray(X, dtype=np.float64).reshape(-1, 3)
    conn = np.asarray(conn, dtype=np.int64).reshape(-1, 8)
    pts = [(0.0, 0.0, 0.0, 8.0)] if reduced else [(p[0] * _GAUSS, p[1] * _GAUSS, p[2] * _GAUSS, 1.0) for p in _HEX_NODES]
    out = np.zeros(3 * len(X))
    xe = X[conn]  # (E, 8, 3)
    for xi, eta, zeta, w in pts:
        f = 1 + _HEX_NODES * np.array([xi, eta, zeta])  # (8, 3)
        N = 0.125 * f[:, 0] * f[:, 1] * f[:, 2]
        dN = 0.125 * np.stack([_HEX_NODES[:, 0] * f[:, 1] * f[:, 2], _HEX_NODES[:, 1] * f[:, 0] * f[:, 2],
                               _HEX_NODES[:, 2] * f[:, 0] * f[:, 1]], axis=1)  # (8, 3)
        J = np.einsum("eai,ak->eik", xe, dN)
        det = np.linalg.det(J)
        for a in range(8):
            np.add.at(out, 3 * conn[:, a] + component, w * N[a] * det)
    return out


def boundary_value(times, values, t):
    times = np.ascontiguousarray(times, dtype=np.float64)
    values = np.ascontiguousarray(values, dtype=np.float64)
    return lib().orc_boundary_value(len(times), _f(times), _f(values), t)


def enforce_dirichlet(rp, ci, val, b, fixed):
    fixed = np.ascontiguousarray(fixed, dtype=np.int32)
    lib().orc_enforce_dirichlet(len(rp) - 1, _i(rp), _i(ci), _f(val), _f(b), len(fixed), _i(fixed))


def spmv(rp, ci, val, x, threads=0):
    y = np.zeros(len(rp) - 1)
    lib().orc_spmv(len(rp) - 1, _i(rp), _i(ci), _f(val), _f(np.ascontiguousarray(x, dtype=np.float64)), _f(y), threads)
    return y


def _krylov(name, rp, ci, val, b, tol, max_it, threads=0):
    rp = np.ascontiguousarray(rp, dtype=np.int32)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(len(rp) - 1)
    its, err = C.c_int32(0), C.c_double(0)
    code = getattr(lib(), name)(len(rp) - 1, _i(rp), _i(ci), _f(val), _f(b), _f(x), tol, max_it, threads,
                                C.byref(its), C.byref(err))
    return code, x, its.value, err.value


def cg_eigen(*a, **k):
    return _krylov("orc_cg_eigen", *a, **k)


def cg_cuda(*a, **k):
    return _krylov("orc_cg_cuda", *a, **k)


def bicgstab_eigen(*a, **k):
    return _krylov("orc_bicgstab_eigen", *a, **k)


def bicgstab_cuda(*a, **k):
    return _krylov("orc_bicgstab_cuda", *a, **k)


def rcm(rp, ci):
    perm = np.zeros(len(rp) - 1, np.int32)
    lib().orc_rcm(len(rp) - 1, _i(rp), _i(ci), _i(perm))
    return perm


class TimeStep:
    def __init__(self, period, initial, minimum, maximum, adaptive, history):
        opt = TimeOptions(period, initial, minimum, maximum, int(adaptive))
        h = np.ascontiguousarray(history, dtype=np.float64)
        self.h = C.c_void_p()
        code = lib().orc_time_create(C.byref(opt), len(h), _f(h), C.byref(self.h))
        if code != OK:
            raise RuntimeError("orc_time_create failed")

    def __del__(self):
        try:
            lib().orc_time_destroy(self.h)
        except Exception:
            pass

    def update(self, converged):
        return lib().orc_time_update(self.h, int(converged))

    def query(self):
        st, lt = C.c_double(), C.c_double()
        fa, su = C.c_int32(), C.c_int32()
        lib().orc_time_query(self.h, C.byref(st), C.byref(lt), C.byref(fa), C.byref(su))
        return dict(step_time=st.value, last_step_time=lt.value, fully_applied=bool(fa.value), successful=su.value)


def newton_options(sim: dict, criterion=STOP_RELATIVE, threads=0, rcm=0, solver=0, record_trajectory=False,
                   latin=None):
    """latin: the search direction of latin_matrix's iterations, or None for static_matrix's."""
    t = sim["time"]
    inc = t["increments"]
    nl = sim["nonlinear_options"]
    ls = sim.get("linear_solver", {})
    return NewtonOptions(
        TimeOptions(t["period"], inc["initial"], inc["minimum"], inc["maximum"], int(bool(inc["adaptive"]))),
        nl["residual_tolerance"], nl["displacement_tolerance"], nl.get("linear_iterations", 10),
        int("absolute_tolerance" in nl), sim.get("norm_mode", 0), ls.get("tolerance", 1e-5),
        ls.get("maximum_iterations", 2000), criterion, solver, threads, rcm, int(record_trajectory),
        int(latin is not None), 0.0 if latin is None else float(latin))
