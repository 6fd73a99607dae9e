"""Result output (neon_amd.output), on the CPU.

* E = N^-1 (shape_function.hpp:70-96 with as many points as nodes) reproduces any trilinear
  field: its values at the Gauss points extrapolate to its nodal values exactly, which fixes
  both the matrix and the point / node orders.
* Node averaging returns a uniform per-point field unchanged and, on an undistorted mesh, the
  nodal values of a field linear in X (each element's trilinear interpolant is exact).
* The VTK files for the reference's J2 cube example solved by the oracle: the collection lists
  the initial and converged states with neon's time format, every .vtu parses, holds the
  requested fields with the right components, the nodal displacement equals the solution and
  the averaged fields equal the homogeneous closed form (1-D return map).
"""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from neon_amd import _capi, ingest, output
from oracle import oracle as orc

import neon_files
import test_ingest as ti


def _trilinear(c, p):
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    return (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * y * z + c[6] * x * z + c[7] * x * y * z)


def test_extrapolation_reproduces_trilinear_fields():
    N, dN = output.shape_at_points()
    E = output.extrapolation_matrix()
    np.testing.assert_allclose(N @ E, np.eye(8), atol=1e-14)
    assert np.allclose(N.sum(axis=1), 1.0) and np.allclose(dN.sum(axis=1), 0.0)
    rng = np.random.default_rng(0)
    for _ in range(5):
        c = rng.normal(size=8)
        at_points = _trilinear(c, output._POINTS)
        at_nodes = _trilinear(c, output._NODES)
        np.testing.assert_allclose(E @ at_points, at_nodes, rtol=0, atol=1e-13)


def test_node_average_uniform_and_linear_fields():
    X, conn, _ = neon_files.cube_mesh(3)
    n = len(X)
    qp = np.tile([[1.5, -2.0]], (8 * len(conn), 1))
    np.testing.assert_allclose(output.node_average(qp, conn, n), np.tile([[1.5, -2.0]], (n, 1)), atol=1e-14)
    # a field linear in X at the Gauss points of each element
    a = np.array([0.3, -1.1, 2.0, 0.7])
    N, _ = output.shape_at_points()
    xq = np.einsum("la,eai->eli", N, X[conn]).reshape(-1, 3)
    f = a[0] + xq @ a[1:]
    np.testing.assert_allclose(output.node_average(f[:, None], conn, n)[:, 0], a[0] + X @ a[1:], atol=1e-13)


def test_deformation_measures_of_a_homogeneous_displacement():
    X, conn, _ = ti._jittered_cube(3, 2)
    G = np.array([[0.1, 0.02, -0.03], [0.0, -0.05, 0.01], [0.04, 0.0, 0.2]])
    H, F = output.deformation_measures(X, conn, (X @ G.T).reshape(-1))
    np.testing.assert_allclose(H, np.broadcast_to(G, H.shape), atol=1e-13)
    np.testing.assert_allclose(F, np.broadcast_to(np.eye(3) + G, F.shape), atol=1e-13)


def test_vtk_files_of_the_j2_cube(tmp_path):
    p = ingest.read(ti.J2_CUBE)
    assert p.output_name == "cube" and "von_mises_stress" in p.visualisation["fields"]
    m = orc.Model(p.X, p.conn, p.params, model=p.model, tangent=orc.TANGENT_REFERENCE)
    code, u, recs = m.newton(orc.newton_options(p.simulation), p.boundaries)
    assert code == orc.OK
    w = output.vtk_file_output(p.output_name, p.visualisation, p.X, p.conn, p.model, str(tmp_path))
    w.write(0, 0.0, w.fields_at(np.zeros(p.n_dofs), np.zeros(p.n_dofs), _zero_state(p)))
    rp, ci = m.pattern()
    _, f = m.assemble(rp, ci)
    w.write(1, 1.0, w.fields_at(u, f, m.qp_state()))
    w.close()
    pvd = ET.parse(tmp_path / "cube.pvd").getroot()
    sets = pvd.find("Collection").findall("DataSet")
    assert [(d.get("timestep"), d.get("file")) for d in sets] == [("0.000000", "visualisation/cube_0.vtu"),
                                                                   ("1.000000", "visualisation/cube_1.vtu")]
    grid = ET.parse(tmp_path / "visualisation" / "cube_1.vtu").getroot()
    piece = grid.find("UnstructuredGrid").find("Piece")
    assert (int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells"))) == (216, 125)
    arrays = {a.get("Name"): a for a in piece.find("PointData").findall("DataArray")}
    assert set(arrays) == set(p.visualisation["fields"])
    comps = {k: int(a.get("NumberOfComponents")) for k, a in arrays.items()}
    assert comps["displacement"] == 3 and comps["cauchy_stress"] == 9 and comps["von_mises_stress"] == 1
    val = {k: np.array(a.text.split(), dtype=float).reshape(216, -1) for k, a in arrays.items()}
    np.testing.assert_array_equal(val["displacement"].reshape(-1), u)
    K, G, sy, H = p.params[0]
    E = 9 * K * G / (3 * K + G)
    nu = (3 * K - 2 * G) / (2 * (3 * K + G))
    sig = (sy + H * 1e-3) * E / (E + H)
    ep = 1e-3 - sig / E
    elat = -nu * sig / E - 0.5 * ep
    np.testing.assert_allclose(val["von_mises_stress"][:, 0], sig, rtol=1e-9)
    np.testing.assert_allclose(val["effective_plastic_strain"][:, 0], ep, rtol=1e-9)
    np.testing.assert_allclose(val["cauchy_stress"][:, 8], sig, rtol=1e-9)
    np.testing.assert_allclose(val["cauchy_stress"][:, [0, 4]], 0.0, atol=1e-9 * sig)
    np.testing.assert_allclose(val["linearised_strain"][:, [0, 4, 8]], np.tile([elat, elat, 1e-3], (216, 1)),
                               rtol=1e-9)
    np.testing.assert_allclose(val["linearised_plastic_strain"][:, 8], ep, rtol=1e-9)
    np.testing.assert_allclose(val["linearised_plastic_strain"][:, 0], -0.5 * ep, rtol=1e-9)
    pts = np.array(piece.find("Points").find("DataArray").text.split(), dtype=float).reshape(-1, 3)
    np.testing.assert_array_equal(pts, p.X)
    cells = {a.get("Name"): np.array(a.text.split(), dtype=int) for a in piece.find("Cells").findall("DataArray")}
    np.testing.assert_array_equal(cells["connectivity"].reshape(-1, 8), p.conn)
    assert set(cells["types"]) == {output.VTK_HEXAHEDRON}


def _zero_state(p):
    Q = 8 * p.n_elem
    z = np.zeros((Q, 3, 3))
    return dict(sigma=z, eps_p=z, detF=np.ones(Q), ebar=np.zeros(Q), vm=np.zeros(Q))


def test_output_names_follow_the_model(tmp_path):
    X, conn, _ = neon_files.cube_mesh(1)
    with pytest.raises(ValueError, match="does not exist in mesh"):
        output.vtk_file_output("a", {"fields": ["von_mises_stress"]}, X, conn, _capi.MODEL_NEOHOOKE, str(tmp_path))
    with pytest.raises(ValueError, match="not a valid output variable"):
        output.vtk_file_output("a", {"fields": ["magnetic_field"]}, X, conn, _capi.MODEL_J2, str(tmp_path))
    with pytest.raises(ValueError, match="visualisation field must be specified"):
        output.vtk_file_output("a", None, X, conn, _capi.MODEL_J2, str(tmp_path))
    assert not os.path.exists(tmp_path / "a.pvd")
