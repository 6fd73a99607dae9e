"""Synthetic hex8 problems: meshes, per-element parameters and Dirichlet boundaries.

Units are MPa and mm-like unit lengths.  Seeds are ``0x6E656F6E`` ("neon") followed by a
two-digit stream id.  Parameters are 4 per element: (bulk K, shear G, yield stress, hardening).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi

SEED_JITTER = 0x6E656F6E01
SEED_MU = 0x6E656F6E02
SEED_LOGMU = 0x6E656F6E03
SEED_TARGET = 0x6E656F6E04
SEED_NODES = 0x6E656F6E05
SEED_DIRECTIONS = 0x6E656F6E06
SEED_YIELD = 0x6E656F6E07


@dataclass
class Problem:
    name: str
    X: np.ndarray           # (N, 3)
    conn: np.ndarray        # (E, 8) int32
    ijk: np.ndarray         # (N, 3) int32 structured indices
    params: np.ndarray      # (E, 4)
    model: int
    boundaries: list        # dicts {name, dofs, times, values}
    simulation: dict
    shape: tuple = ()
    extra: dict = field(default_factory=dict)

    @property
    def n_nodes(self):
        return self.X.shape[0]

    @property
    def n_elem(self):
        return self.conn.shape[0]

    @property
    def n_dofs(self):
        return 3 * self.X.shape[0]

    def fixed_dofs(self) -> np.ndarray:
        return np.unique(np.concatenate([b["dofs"] for b in self.boundaries])).astype(np.int32)


def box_mesh(nx, ny, nz, L=(1.0, 1.0, 1.0), jitter=0.0, seed=SEED_JITTER, hole_radius=0.0):
    lib = _capi.lib()
    nn, ne = C.c_int32(0), C.c_int32(0)
    code = lib.neon_mesh_box_size(nx, ny, nz, L[0], L[1], L[2], hole_radius, C.byref(nn), C.byref(ne))
    if code != 0:
        raise _capi.NeonError(code, "neon_mesh_box_size")
    X = np.zeros((nn.value, 3))
    conn = np.zeros((ne.value, 8), np.int32)
    ijk = np.zeros((nn.value, 3), np.int32)
    code = lib.neon_mesh_box(nx, ny, nz, L[0], L[1], L[2], jitter, seed, hole_radius, _capi.f64(X), _capi.i32(conn),
                             _capi.i32(ijk))
    if code != 0:
        raise _capi.NeonError(code, "neon_mesh_box")
    return X, conn, ijk


def uniform(seed, n, lo, hi):
    out = np.zeros(n)
    _capi.lib().neon_random_uniform(seed, n, lo, hi, _capi.f64(out))
    return out


def normal(seed, n):
    out = np.zeros(n)
    _capi.lib().neon_random_normal(seed, n, _capi.f64(out))
    return out


def face_dofs(ijk, axis, index, component):
    nodes = np.nonzero(ijk[:, axis] == index)[0]
    return (3 * nodes + component).astype(np.int32)


def simulation(period=1.0, steps=5, residual_tolerance=1e-8, displacement_tolerance=1e-8, solver_tolerance=1e-10,
               max_iterations=20000, precision="fp64", newton_iterations=10, norm_mode=0, max_load_steps=0):
    inc = period / steps
    return {
        "time": {"period": period, "increments": {"initial": inc, "minimum": inc, "maximum": inc, "adaptive": False}},
        "nonlinear_options": {"displacement_tolerance": displacement_tolerance,
                              "residual_tolerance": residual_tolerance, "linear_iterations": newton_iterations},
        "linear_solver": {"type": "iterative", "device": "gpu", "tolerance": solver_tolerance,
                          "maximum_iterations": max_iterations, "backend": {"type": "hip", "precision": precision}},
        "norm_mode": norm_mode,
        "max_load_steps": max_load_steps,
    }


def cube_clamped(n=10, mu=0.4, bulk=2.0, top=0.01, steps=5, **sim):
    """Unit cube n^3, z=0 clamped, top u_z = +top over t in [0,1] (the 10^3 CPU case at n=10)."""
    X, conn, ijk = box_mesh(n, n, n)
    E = conn.shape[0]
    params = np.zeros((E, 4))
    params[:, 0] = bulk
    params[:, 1] = mu
    bcs = [
        {"name": "bottom_x", "dofs": face_dofs(ijk, 2, 0, 0), "times": [0.0, 1.0], "values": [0.0, 0.0]},
        {"name": "bottom_y", "dofs": face_dofs(ijk, 2, 0, 1), "times": [0.0, 1.0], "values": [0.0, 0.0]},
        {"name": "bottom_z", "dofs": face_dofs(ijk, 2, 0, 2), "times": [0.0, 1.0], "values": [0.0, 0.0]},
        {"name": "top_z", "dofs": face_dofs(ijk, 2, n, 2), "times": [0.0, 1.0], "values": [0.0, top]},
    ]
    return Problem(f"cube{n}", X, conn, ijk, params, _capi.MODEL_NEOHOOKE, bcs,
                   simulation(steps=steps, **sim), (n, n, n))


def block_rollers(n, stretch=0.2, steps=10, jitter=0.1, mu_seed=SEED_MU, dims=None, **sim):
    """Jittered n^3 neo-Hooke block, mu_e ~ U[0.5, 1.5] MPa, K = 10 mu; rollers on x=0, y=0,
    z=0 and top u_z = stretch over 10 equal steps (the 64^3 / 128^3 cases).  dims = (nx, ny,
    nz) cells gives a box of unit cells instead (top at z = nz)."""
    nx, ny, nz = dims if dims is not None else (n, n, n)
    X, conn, ijk = box_mesh(nx, ny, nz, L=(nx / n, ny / n, nz / n), jitter=jitter, seed=SEED_JITTER)
    E = conn.shape[0]
    mu = uniform(mu_seed, E, 0.5, 1.5)
    params = np.zeros((E, 4))
    params[:, 0] = 10.0 * mu
    params[:, 1] = mu
    bcs = [
        {"name": "x0", "dofs": face_dofs(ijk, 0, 0, 0), "times": [0.0, 1.0], "values": [0.0, 0.0]},
        {"name": "y0", "dofs": face_dofs(ijk, 1, 0, 1), "times": [0.0, 1.0], "values": [0.0, 0.0]},
        {"name": "z0", "dofs": face_dofs(ijk, 2, 0, 2), "times": [0.0, 1.0], "values": [0.0, 0.0]},
        {"name": "top", "dofs": face_dofs(ijk, 2, nz, 2), "times": [0.0, 1.0], "values": [0.0, stretch]},
    ]
    name = f"block{n}" if dims is None else f"block{nx}x{ny}x{nz}"
    return Problem(name, X, conn, ijk, params, _capi.MODEL_NEOHOOKE, bcs, simulation(steps=steps, **sim),
                   (nx, ny, nz))


def notched_plate_j2(nx=128, ny=64, nz=64, strain=0.005, steps=20, E_mod=200.0e3, nu=0.3, H=2000.0,
                     sy_range=(225.0, 275.0), hole_cells=8.0, jitter=0.0, **sim):
    """Small-strain J2 plate 2 x 1 x 1 (nx x ny x nz cells), central z-parallel hole of radius
    hole_cells * h, per-element yield stress ~ U[225, 275] MPa; x=0 roller, x=2 face pulled to
    `strain` nominal strain, rollers on y=0 and z=0."""
    h = 1.0 / ny
    L = (nx * h, ny * h, nz * h)
    X, conn, ijk = box_mesh(nx, ny, nz, L, jitter=jitter, hole_radius=hole_cells * h)
    E = conn.shape[0]
    bulk = E_mod / (3.0 * (1.0 - 2.0 * nu))
    shear = E_mod / (2.0 * (1.0 + nu))
    params = np.zeros((E, 4))
    params[:, 0] = bulk
    params[:, 1] = shear
    params[:, 2] = uniform(SEED_YIELD, E, *sy_range)
    params[:, 3] = H
    bcs = [
        {"name": "x0", "dofs": face_dofs(ijk, 0, 0, 0), "times": [0.0, 1.0], "values": [0.0, 0.0]},
        {"name": "y0", "dofs": face_dofs(ijk, 1, 0, 1), "times": [0.0, 1.0], "values": [0.0, 0.0]},
        {"name": "z0", "dofs": face_dofs(ijk, 2, 0, 2), "times": [0.0, 1.0], "values": [0.0, 0.0]},
        {"name": "x1", "dofs": face_dofs(ijk, 0, nx, 0), "times": [0.0, 1.0], "values": [0.0, strain * L[0]]},
    ]
    return Problem(f"plate{nx}x{ny}x{nz}", X, conn, ijk, params, _capi.MODEL_J2, bcs,
                   simulation(steps=steps, **sim), (nx, ny, nz))


DAMAGE_CONSTANTS = (5500.0, 250.0, 1220.0, 2.5, 0.6, 2.0)


def damage_cube(n=1, top=0.018, period=0.01, steps=10, E_mod=134.0e3, nu=0.3, yield_stress=85.0,
                newton_iterations=30, **sim):
    """The viscoplastic damage cube of neon's examples (examples/solid_mechanics/viscoplastic_damage/
    cube/cube.json): a unit cube (n^3 elements; the example has one), symmetry rollers on x=0,
    y=0, z=0, top u_z = `top` over t in [0, period] in `steps` fixed increments, the example's
    steel (E = 134 GPa in MPa units, nu = 0.3, yield 85 MPa) with the damage constants
    DAMAGE_CONSTANTS (kinematic hardening modulus 5500, softening multiplier 250, plasticity
    viscous denominator 1220 and exponent 2.5, damage viscous denominator 0.6 and exponent 2),
    BiCGSTAB (the model's tangent is unsymmetric)."""
    X, conn, ijk = box_mesh(n, n, n)
    E = conn.shape[0]
    params = np.zeros((E, 4))
    params[:, 0] = E_mod / (3.0 * (1.0 - 2.0 * nu))
    params[:, 1] = E_mod / (2.0 * (1.0 + nu))
    params[:, 2] = yield_stress
    bcs = [
        {"name": "Xsym", "dofs": face_dofs(ijk, 0, 0, 0), "times": [0.0, period], "values": [0.0, 0.0]},
        {"name": "Ysym", "dofs": face_dofs(ijk, 1, 0, 1), "times": [0.0, period], "values": [0.0, 0.0]},
        {"name": "Zsym", "dofs": face_dofs(ijk, 2, 0, 2), "times": [0.0, period], "values": [0.0, 0.0]},
        {"name": "ZLoad", "dofs": face_dofs(ijk, 2, n, 2), "times": [0.0, period], "values": [0.0, top]},
    ]
    simu = simulation(period=period, steps=steps, newton_iterations=newton_iterations, **sim)
    simu["damage_constants"] = list(DAMAGE_CONSTANTS)
    return Problem(f"damage_cube{n}", X, conn, ijk, params, _capi.MODEL_J2_DAMAGE, bcs, simu, (n, n, n))


def smoothed_log_mu(n, seed, sigma=0.3, passes=3):
    """log mu_e ~ N(0, sigma^2) on an n^3 element grid, smoothed by `passes` 27-point
    (3x3x3, edge-clamped) averages; returns mu_e in element order."""
    g = sigma * normal(seed, n ** 3).reshape(n, n, n)  # [k, j, i]
    for _ in range(passes):
        p = np.pad(g, 1, mode="edge")
        acc = np.zeros_like(g)
        for dk in range(3):
            for dj in range(3):
                for di in range(3):
                    acc += p[dk:dk + n, dj:dj + n, di:di + n]
        g = acc / 27.0
    return np.exp(g.reshape(-1))


def adjoint_block(n=96, stretch=0.05, steps=1, n_obs=4096, **sim):
    """Neo-Hooke n^3 block with smoothed log-normal mu (K = 10 mu), rollers on x=0, y=0, z=0
    and top u_z = stretch; observations at n_obs seeded top-surface nodes against a target
    displacement field solved with a different seeded mu."""
    X, conn, ijk = box_mesh(n, n, n, jitter=0.1, seed=SEED_JITTER)
    mu = smoothed_log_mu(n, SEED_LOGMU)
    mu_target = smoothed_log_mu(n, SEED_TARGET)
    params = np.zeros((conn.shape[0], 4))
    params[:, 0] = 10.0 * mu
    params[:, 1] = mu
    top = np.nonzero(ijk[:, 2] == n)[0]
    keys = uniform(SEED_NODES, len(top), 0.0, 1.0)
    obs = np.sort(top[np.argsort(keys, kind="stable")[: min(n_obs, len(top))]])
    bcs = [
        {"name": "x0", "dofs": face_dofs(ijk, 0, 0, 0), "times": [0.0, 1.0], "values": [0.0, 0.0]},
        {"name": "y0", "dofs": face_dofs(ijk, 1, 0, 1), "times": [0.0, 1.0], "values": [0.0, 0.0]},
        {"name": "z0", "dofs": face_dofs(ijk, 2, 0, 2), "times": [0.0, 1.0], "values": [0.0, 0.0]},
        {"name": "top", "dofs": face_dofs(ijk, 2, n, 2), "times": [0.0, 1.0], "values": [0.0, stretch]},
    ]
    return Problem(f"adjoint{n}", X, conn, ijk, params, _capi.MODEL_NEOHOOKE, bcs,
                   simulation(steps=steps, **sim), (n, n, n),
                   extra={"mu": mu, "mu_target": mu_target, "observed_nodes": obs})
