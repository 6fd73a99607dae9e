"""Result output in neon's layout: a ``<name>.pvd`` collection beside the input and one VTK
XML unstructured grid ``visualisation/<name>_<step>.vtu`` per write (vtk_file_output,
src/io/file_output.cpp:66-162), written where neon's static_matrix calls mesh.write (the
initial state, after a cut-back, after each converged step; ``static_matrix.solve``'s
``on_step`` hook).

Fields (mesh.cpp:174-214, variable_string_adapter.cpp): the nodal ``displacement`` and
``reaction_force`` (minus the internal force of the last assembly, mesh.hpp:44), and internal
variables extrapolated from the 8 quadrature points to the element's nodes with E = N^-1
(N_lj = N_j at point l; shape_function.hpp:70-96 with m = n) and averaged without weights over
the elements sharing a node (node_averaged_variables.hpp:15-40, submesh.cpp:271-344):
``cauchy_stress``, ``deformation_gradient``, ``displacement_gradient`` (9 components, row
major), ``DetF`` for every model; ``linearised_strain``, ``linearised_plastic_strain``,
``von_mises_stress``, ``effective_plastic_strain`` for small-strain J2; ``linearised_strain``
and ``von_mises_stress`` for linear elasticity.  The deformation
measures are recomputed on the host from u as submesh::update_deformation_measures does
(submesh.cpp:210-237: H = (x - X) dN (X dN)^-1, F = (x dN)(X dN)^-1).

Differences from neon: the files are ASCII VTK (neon's default is VTK's binary encoding; the
values are the same), and the collection is closed by ``close()`` rather than a destructor.
"""
from __future__ import annotations

import os
from xml.sax.saxutils import quoteattr

import numpy as np

from . import _capi

VTK_HEXAHEDRON = 12
_GAUSS = 1.0 / np.sqrt(3.0)
# local node coordinates (hexahedron.cpp:21-28) and the 8 quadrature points in the order of
# the device's per-point state (hexahedron_quadrature.cpp:30-38)
_NODES = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                   [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)
_POINTS = _GAUSS * _NODES

NODAL = ("displacement", "reaction_force")
COMMON = ("cauchy_stress", "deformation_gradient", "displacement_gradient", "DetF")
J2_ONLY = ("linearised_strain", "linearised_plastic_strain", "von_mises_stress", "effective_plastic_strain")
DAMAGE_ONLY = ("back_stress", "kinematic_hardening", "damage", "energy_release_rate")
# the variable names each model adds (constitutive_model::variable_names): small-strain J2
# (small_strain_J2_plasticity.cpp:14-26 plus those of isotropic_linear_elasticity.cpp:15-19), J2
# with damage (those and small_strain_J2_plasticity_damage.cpp:21-29), linear elasticity
# (isotropic_linear_elasticity.cpp:15-19); the microsphere models add none
MODEL_VARIABLES = {_capi.MODEL_J2: J2_ONLY, _capi.MODEL_J2_DAMAGE: J2_ONLY + DAMAGE_ONLY,
                   _capi.MODEL_LINEAR_ELASTIC: ("linearised_strain", "von_mises_stress")}


def shape_at_points():
    """N (8 points x 8 nodes) and dN (8 points, 8 nodes, 3) of hexahedron8 at the Gauss points."""
    f = 1.0 + _POINTS[:, None, :] * _NODES[None, :, :]  # (l, a, 3)
    N = 0.125 * f[:, :, 0] * f[:, :, 1] * f[:, :, 2]
    dN = 0.125 * np.stack([_NODES[None, :, 0] * f[:, :, 1] * f[:, :, 2],
                           _NODES[None, :, 1] * f[:, :, 0] * f[:, :, 2],
                           _NODES[None, :, 2] * f[:, :, 0] * f[:, :, 1]], axis=2)
    return N, dN


def extrapolation_matrix() -> np.ndarray:
    """E = N^-1: quadrature-point values to the element's nodes."""
    N, _ = shape_at_points()
    return np.linalg.inv(N)


def node_average(qp_values: np.ndarray, conn: np.ndarray, n_nodes: int) -> np.ndarray:
    """(E*8, c) per-point values -> (n_nodes, c): extrapolated per element, then the unweighted
    average over the elements sharing each node."""
    conn = np.asarray(conn).reshape(-1, 8)
    v = np.asarray(qp_values, dtype=np.float64).reshape(len(conn), 8, -1)
    nodal = np.einsum("al,elc->eac", extrapolation_matrix(), v)
    total = np.zeros((n_nodes, v.shape[2]))
    count = np.zeros(n_nodes)
    np.add.at(total, conn.reshape(-1), nodal.reshape(-1, v.shape[2]))
    np.add.at(count, conn.reshape(-1), 1.0)
    count[count == 0] = 1.0
    return total / count[:, None]


def deformation_measures(X: np.ndarray, conn: np.ndarray, u: np.ndarray):
    """Per point (E*8, 3, 3): displacement gradient H and deformation gradient F."""
    conn = np.asarray(conn).reshape(-1, 8)
    X = np.asarray(X).reshape(-1, 3)
    x = X + np.asarray(u).reshape(-1, 3)
    _, dN = shape_at_points()
    Xe, xe = X[conn], x[conn]  # (E, 8, 3)
    F0 = np.einsum("eai,lak->elik", Xe, dN)
    Fxi = np.einsum("eai,lak->elik", xe, dN)
    F0i = np.linalg.inv(F0)
    H = np.einsum("elij,eljk->elik", Fxi - F0, F0i)
    F = np.einsum("elij,eljk->elik", Fxi, F0i)
    return H.reshape(-1, 3, 3), F.reshape(-1, 3, 3)


def _array(name, data, components, fmt="%.17g"):
    body = " ".join(fmt % v for v in np.asarray(data).reshape(-1))
    ncomp = f' NumberOfComponents="{components}"' if components else ""
    nm = f" Name={quoteattr(name)}" if name else ""
    dtype = "Float64" if fmt != "%d" else "Int64"
    return f'        <DataArray type="{dtype}"{nm}{ncomp} format="ascii">{body}</DataArray>\n'


class vtk_file_output:
    """neon's vtk_file_output for hex8 meshes (io/file_output.cpp:28-162)."""

    def __init__(self, file_name: str, visualisation: dict | None, X, conn, model: int, root: str = "."):
        if visualisation is None:
            raise ValueError("visualisation field must be specified")
        if not file_name:
            raise ValueError('"name" field for file IO must be specified')
        self.fields = list(dict.fromkeys(visualisation.get("fields", [])))
        allowed = NODAL + COMMON + MODEL_VARIABLES.get(model & 0xFF, ())
        for f in self.fields:
            if f not in NODAL + COMMON + J2_ONLY + DAMAGE_ONLY:
                raise ValueError(f"Name {f} is not a valid output variable name")
            if f not in allowed:
                raise ValueError(f"Internal variable {f} does not exist in mesh")
        self.name, self.root, self.model = file_name, root, model
        self.X = np.asarray(X, dtype=np.float64).reshape(-1, 3)
        self.conn = np.asarray(conn).reshape(-1, 8)
        self.directory = "visualisation"
        os.makedirs(os.path.join(root, self.directory), exist_ok=True)
        self.pvd = open(os.path.join(root, file_name + ".pvd"), "w")
        self.pvd.write('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1">\n  <Collection>\n')
        self.written: list[str] = []

    def fields_at(self, u: np.ndarray, f_int: np.ndarray, qp: dict) -> dict:
        """Nodal arrays of the requested fields."""
        n = len(self.X)
        out = {}
        need_measures = any(f in self.fields for f in ("deformation_gradient", "displacement_gradient",
                                                       "linearised_strain"))
        H = F = None
        if need_measures:
            H, F = deformation_measures(self.X, self.conn, u)
        for f in self.fields:
            if f == "displacement":
                out[f] = np.asarray(u).reshape(n, 3)
            elif f == "reaction_force":
                out[f] = -np.asarray(f_int).reshape(n, 3)
            elif f == "cauchy_stress":
                out[f] = node_average(qp["sigma"].reshape(-1, 9), self.conn, n)
            elif f == "deformation_gradient":
                out[f] = node_average(F.reshape(-1, 9), self.conn, n)
            elif f == "displacement_gradient":
                out[f] = node_average(H.reshape(-1, 9), self.conn, n)
            elif f == "DetF":
                out[f] = node_average(qp["detF"].reshape(-1, 1), self.conn, n)
            elif f == "linearised_strain":
                out[f] = node_average((0.5 * (H + np.swapaxes(H, 1, 2))).reshape(-1, 9), self.conn, n)
            elif f == "linearised_plastic_strain":
                out[f] = node_average(qp["eps_p"].reshape(-1, 9), self.conn, n)
            elif f == "von_mises_stress":
                out[f] = node_average(qp["vm"].reshape(-1, 1), self.conn, n)
            elif f == "effective_plastic_strain":
                out[f] = node_average(qp["ebar"].reshape(-1, 1), self.conn, n)
            elif f == "back_stress":
                out[f] = node_average(qp["back_stress"].reshape(-1, 9), self.conn, n)
            elif f == "kinematic_hardening":
                # neon's J2-damage model allocates this tensor but never writes it
                # (perform_radial_return takes it without using it,
                # small_strain_J2_plasticity_damage.cpp:113-115): zero at every write
                out[f] = np.zeros((n, 9))
            elif f == "damage":
                out[f] = node_average(qp["damage"].reshape(-1, 1), self.conn, n)
            elif f == "energy_release_rate":
                out[f] = node_average(qp["energy"].reshape(-1, 1), self.conn, n)
        return out

    def write(self, step: int, time: float, fields: dict) -> str:
        rel = f"{self.directory}/{self.name}_{step}.vtu"
        n, e = len(self.X), len(self.conn)
        parts = ['<?xml version="1.0"?>\n',
                 '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n',
                 "  <UnstructuredGrid>\n", f'    <Piece NumberOfPoints="{n}" NumberOfCells="{e}">\n',
                 "      <PointData>\n"]
        for name, arr in fields.items():
            arr = np.asarray(arr).reshape(n, -1)
            parts.append("  " + _array(name, arr, arr.shape[1]))
        parts += ["      </PointData>\n", "      <Points>\n", "  " + _array(None, self.X, 3), "      </Points>\n",
                  "      <Cells>\n",
                  "  " + _array("connectivity", self.conn, None, "%d"),
                  "  " + _array("offsets", 8 * np.arange(1, e + 1), None, "%d"),
                  f'          <DataArray type="UInt8" Name="types" format="ascii">'
                  f'{" ".join([str(VTK_HEXAHEDRON)] * e)}</DataArray>\n',
                  "      </Cells>\n", "    </Piece>\n", "  </UnstructuredGrid>\n", "</VTKFile>\n"]
        with open(os.path.join(self.root, rel), "w") as fh:
            fh.write("".join(parts))
        # neon prints the time with std::to_string (six decimals)
        self.pvd.write(f'    <DataSet timestep = "{time:.6f}" file = "{rel}" />\n')
        self.pvd.flush()
        self.written.append(rel)
        return rel

    def close(self) -> None:
        if self.pvd and not self.pvd.closed:
            self.pvd.write("  </Collection>\n</VTKFile>\n")
            self.pvd.close()


def solve_with_output(problem, ctx, root: str = ".", precision: str = "fp64", tangent: int | None = None):
    """Run an ingested problem on the device, writing neon's output files under `root` when the
    problem requests visualisation.  Returns (u, static_matrix, writer or None)."""
    import neon_amd

    tangent = neon_amd.TANGENT_REFERENCE if tangent is None else tangent
    mesh = neon_amd.DeviceMesh(problem.X, problem.conn, problem.params, model=problem.model, tangent=tangent,
                               context=ctx, ageing=problem.ageing, damage=getattr(problem, "damage", None))
    sim = dict(problem.simulation)
    sim["linear_solver"] = dict(sim["linear_solver"], backend={"type": "hip", "precision": precision})
    driver = neon_amd.latin_matrix if problem.solution == "latin" else neon_amd.static_matrix
    sm = driver(mesh, sim, problem.boundaries, loads=problem.loads)
    writer = None
    if problem.visualisation is not None:
        writer = vtk_file_output(problem.output_name, problem.visualisation, problem.X, problem.conn, problem.model,
                                 root)

        def on_step(step, time):
            _, f = mesh.download(stiffness=False)
            qp = mesh.qp_state()
            if problem.model == neon_amd.MODEL_J2_DAMAGE:
                qp.update(mesh.damage_state())
            writer.write(step, time, writer.fields_at(mesh.displacement(), f, qp))
    try:
        u = sm.solve(on_step=on_step if writer else None)
    finally:
        if writer:
            writer.close()
    return u, sm, writer
