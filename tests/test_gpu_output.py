"""Result output through the device Newton's step hook (static_matrix.solve(on_step=...)):
the reference's neo-Hooke cube example written as neon writes it, and the hook's contract."""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import neon_amd
from neon_amd import ingest, output

import test_ingest as ti

pytestmark = pytest.mark.gpu


def _read(path):
    piece = ET.parse(path).getroot().find("UnstructuredGrid").find("Piece")
    n = int(piece.get("NumberOfPoints"))
    return {a.get("Name"): np.array(a.text.split(), dtype=float).reshape(n, -1)
            for a in piece.find("PointData").findall("DataArray")}


def test_neohooke_cube_example_output(ctx, tmp_path):
    p = ingest.read(ti.NEOHOOKE_CUBE)
    u, sm, w = output.solve_with_output(p, ctx, root=str(tmp_path))
    sets = ET.parse(tmp_path / "cube.pvd").getroot().find("Collection").findall("DataSet")
    # the initial state (step 0, time 0) and the converged step (the whole load in one increment)
    assert [(d.get("timestep"), d.get("file")) for d in sets] == [("0.000000", "visualisation/cube_0.vtu"),
                                                                   ("1.000000", "visualisation/cube_1.vtu")]
    first, last = _read(tmp_path / "visualisation" / "cube_0.vtu"), _read(tmp_path / "visualisation" / "cube_1.vtu")
    assert set(last) == {"displacement", "deformation_gradient", "cauchy_stress"}
    np.testing.assert_array_equal(last["displacement"].reshape(-1), u)
    # homogeneous uniaxial stretch 1.5: F = diag(l, l, 1.5) and the Cauchy stress
    # sigma_zz = (lambda ln J + mu (1.5^2 - 1)) / J, no lateral stress; the example's own Newton
    # tolerance (1e-3) leaves ~1e-6 relative
    K, G = p.params[0, :2]
    lam = K - 2.0 * G / 3.0
    ll = ti.uniaxial_neohooke_lateral(K, G, 1.5)
    J = ll * ll * 1.5
    szz = (lam * np.log(J) + G * (1.5 ** 2 - 1.0)) / J
    np.testing.assert_allclose(last["deformation_gradient"][:, [0, 4, 8]], np.tile([ll, ll, 1.5], (len(u) // 3, 1)),
                               rtol=1e-5)
    np.testing.assert_allclose(last["cauchy_stress"][:, 8], szz, rtol=1e-4)
    assert np.abs(last["cauchy_stress"][:, [0, 4]]).max() < 1e-4 * szz
    # the initial state: no displacement, and a stress at roundoff (F = (x dN)(X dN)^-1 is the
    # identity to ~1e-16, times lambda ~ 1.6e8 for this nearly incompressible rubber)
    assert not first["displacement"].any()
    assert np.abs(first["cauchy_stress"]).max() < 1e-12 * szz


def test_step_hook_sees_every_converged_state(ctx, tmp_path):
    # four fixed increments: the hook runs at the initial state and after each step, and an
    # exception raised inside it reaches the caller after the solve
    p = ti._traction_problem(tmp_path, "traction", 0.4)
    mesh = neon_amd.DeviceMesh(p.X, p.conn, p.params, model=p.model, context=ctx)
    sm = neon_amd.static_matrix(mesh, p.simulation, p.boundaries, loads=p.loads)
    seen = []
    u = sm.solve(on_step=lambda step, time: seen.append((step, time, np.abs(mesh.displacement()).max())))
    assert [(s, t) for s, t, _ in seen] == [(0, 0.0), (1, 0.25), (2, 0.5), (3, 0.75), (4, 1.0)]
    assert seen[0][2] == 0.0 and seen[-1][2] == np.abs(u).max()
    assert all(a[2] < b[2] for a, b in zip(seen, seen[1:]))

    def fail(step, time):
        if step == 2:
            raise RuntimeError("stop here")

    with pytest.raises(RuntimeError, match="stop here"):
        sm.solve(on_step=fail)
