"""Writer of small neon input files (simulation JSON + ``.mesh``) for the ingest tests: an
n^3 unit cube of hex8 cells with quadrilateral face groups x0, x1, y0, y1, z0, z1 (outward
normals, counter-clockwise seen from outside) and the volume group named like the part."""
from __future__ import annotations

import json
import os

import numpy as np

GMSH_QUAD4 = 3
GMSH_HEX8 = 5


def cube_mesh(n: int):
    """Coordinates (N, 3), hex8 connectivity and the six face groups of the unit cube."""
    m = n + 1
    ii, jj, kk = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    node = lambda i, j, k: i + m * (j + m * k)  # noqa: E731
    X = np.zeros((m ** 3, 3))
    X[node(ii, jj, kk).ravel()] = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1) / n
    conn = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                conn.append([node(i, j, k), node(i + 1, j, k), node(i + 1, j + 1, k), node(i, j + 1, k),
                             node(i, j, k + 1), node(i + 1, j, k + 1), node(i + 1, j + 1, k + 1), node(i, j + 1, k + 1)])
    faces = {"x0": [], "x1": [], "y0": [], "y1": [], "z0": [], "z1": []}
    for a in range(n):
        for b in range(n):
            faces["z1"].append([node(a, b, n), node(a + 1, b, n), node(a + 1, b + 1, n), node(a, b + 1, n)])
            faces["z0"].append([node(a, b, 0), node(a, b + 1, 0), node(a + 1, b + 1, 0), node(a + 1, b, 0)])
            faces["x1"].append([node(n, a, b), node(n, a + 1, b), node(n, a + 1, b + 1), node(n, a, b + 1)])
            faces["x0"].append([node(0, a, b), node(0, a, b + 1), node(0, a + 1, b + 1), node(0, a + 1, b)])
            faces["y1"].append([node(a, n, b), node(a, n, b + 1), node(a + 1, n, b + 1), node(a + 1, n, b)])
            faces["y0"].append([node(a, 0, b), node(a + 1, 0, b), node(a + 1, 0, b + 1), node(a, 0, b + 1)])
    return X, np.array(conn), {k: np.array(v) for k, v in faces.items()}


def write_cube_simulation(directory, n, material, constitutive, boundaries, time=None, nonlinear=None,
                          linear_solver=None, part="cube", extra_groups=None):
    """Write ``<part>.json`` and ``<part>.mesh`` for the n^3 cube into `directory`; returns the
    JSON path.  `extra_groups`: more element groups {name: (gmsh type, connectivity)}."""
    X, conn, faces = cube_mesh(n)
    groups = [{"Name": name, "Type": GMSH_QUAD4, "NodalConnectivity": f.tolist()} for name, f in faces.items()]
    groups.append({"Name": part, "Type": GMSH_HEX8, "NodalConnectivity": conn.tolist()})
    for name, (gtype, c) in (extra_groups or {}).items():
        groups.append({"Name": name, "Type": gtype, "NodalConnectivity": np.asarray(c).tolist()})
    mesh = {"Elements": groups, "Nodes": [{"Coordinates": X.tolist(), "Indices": list(range(len(X)))}]}
    sim = {
        "name": part,
        "parts": [{"name": part, "material": "mat"}],
        "materials": [dict(material, name="mat")],
        "steps": [{
            "name": "load",
            "module": "solid_mechanics",
            "solution": "equilibrium",
            "nonlinear_options": nonlinear or {"residual_tolerance": 1e-10, "displacement_tolerance": 1e-10},
            "time": time or {"period": 1.0, "increments": {"initial": 0.25, "minimum": 0.001, "maximum": 0.25,
                                                           "adaptive": False}},
            "meshes": [{"name": part, "constitutive": constitutive, "element_options": {"quadrature": "full"},
                        "boundaries": boundaries}],
            "linear_solver": linear_solver or {"type": "iterative", "tolerance": 1e-12, "maximum_iterations": 20000},
        }],
    }
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, part + ".mesh"), "w") as fh:
        json.dump(mesh, fh)
    path = os.path.join(directory, part + ".json")
    with open(path, "w") as fh:
        json.dump(sim, fh)
    return path


def rollers(extra):
    """Symmetry rollers on x0, y0, z0 plus `extra` boundaries."""
    return [
        {"name": "x0", "type": "displacement", "time": [0.0, 1.0], "x": [0.0, 0.0]},
        {"name": "y0", "type": "displacement", "time": [0.0, 1.0], "y": [0.0, 0.0]},
        {"name": "z0", "type": "displacement", "time": [0.0, 1.0], "z": [0.0, 0.0]},
    ] + list(extra)
