"""Reader of neon's input files for the device path.

A neon simulation is a JSON file (src/simulation_parser.cpp:45-145) naming its parts, their
materials and the load steps; each part's mesh is ``<part name>.mesh`` beside it, a JSON file
with ``Nodes[0].Coordinates`` (nodal_coordinates.cpp:13-33) and ``Elements``: groups with a
``Name``, a gmsh element ``Type`` and a ``NodalConnectivity`` list (basic_mesh.cpp:8-29,
basic_submesh.cpp:12-52).  The part's volume elements are the groups named like the part;
boundary groups carry the boundary names.  Hexahedron8 (gmsh 5) and quadrilateral4 (gmsh 3)
keep gmsh's node order (node_ordering_adapter.cpp:46-136 reorders neither).

``read(path)`` returns a :class:`Problem` for :class:`neon_amd.DeviceMesh` and
:class:`neon_amd.static_matrix`:

* materials (isotropic_elastic_property.cpp:10-43): bulk and shear moduli, or the elastic
  modulus and Poisson's ratio (ratio < 0.5); J2 also ``yield_stress`` and the optional
  ``isotropic_hardening_modulus`` (isotropic_elastic_plastic.cpp:10-36);
* constitutive models (constitutive_model_factory.cpp:37-160): ``neohooke`` /
  ``compressible_neohooke``, ``isotropic_linear_elasticity``, ``J2_plasticity`` with
  ``finite_strain: false``, and ``microsphere`` (affine with gaussian or langevin statistics,
  nonaffine; quadrature BO21 or BO33) with ``segments_per_chain`` (and
  ``nonaffine_stretch_parameter``) in the material (micromechanical_elastomer.cpp:12-20);
* boundaries (mesh.cpp:85-143, nonfollower_load.cpp:12-122): ``displacement`` per component
  key x / y / z on the unique nodes of the named groups; ``traction`` (per component) and
  ``pressure`` (``value``) on quadrilateral faces and ``body_force`` (per component) on hex8
  groups, as unit-load vectors on the initial configuration times the history; and
  ``nodal_force`` (per component, 1 at each unique node, applied at its t = 1 value in every
  step as neon's compute_external_force does);
* the ``linear_solver`` block: ``iterative`` keeps its tolerance and iteration limit (neon's
  defaults 1e-5 and 2000, linear_solver.hpp:54-55); the direct solvers (``PaStiX``,
  ``MUMPS``, ``direct``) have no device counterpart and become CG to a relative 1e-12.

Everything else (multibody steps, other element types, generated sinusoidal histories,
other constitutive models) raises ``ValueError`` naming what is not supported.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field

import numpy as np

from . import _capi

GMSH_QUAD4 = 3
GMSH_HEX8 = 5
DOF_TABLE = {"x": 0, "y": 1, "z": 2}
DIRECT_SOLVERS = ("PaStiX", "MUMPS", "direct")
DIRECT_TOLERANCE = 1.0e-12
DIRECT_MAX_ITERATIONS = 100000


@dataclass
class Problem:
    name: str
    X: np.ndarray          # (N, 3) coordinates
    conn: np.ndarray       # (E, 8) hex8 connectivity
    params: np.ndarray     # (E, 4): bulk, shear, yield stress, hardening
    model: int             # MODEL_* (with SPHERE_BO33 for the 33-direction microsphere rule)
    boundaries: list       # Dirichlet: {name, dofs, times, values}
    loads: list            # nonfollower: {name, dofs, values, times, loads[, fixed_time]}
    simulation: dict       # static_matrix options
    groups: dict = field(default_factory=dict)  # name -> list of (gmsh type, connectivity array)
    visualisation: dict | None = None  # the step mesh's "visualisation" block (output fields)
    output_name: str = ""              # file name of the output (the step mesh's "name")
    solution: str = "equilibrium"      # the step's "solution": "equilibrium", "latin" or "natural_frequency"
    eigen_solver: dict | None = None   # the step's "eigen_solver" block (natural frequency)
    density: np.ndarray | None = None  # (E,) the material's "density", when given
    ageing: tuple | None = None        # ageing microsphere: (scission, recombination) probabilities
    damage: tuple | None = None        # J2 with damage: the six constants of neon_mesh_set_damage

    @property
    def n_nodes(self) -> int:
        return len(self.X)

    @property
    def n_elem(self) -> int:
        return len(self.conn)

    @property
    def n_dofs(self) -> int:
        return 3 * len(self.X)

    def fixed_dofs(self) -> np.ndarray:
        d = [np.asarray(b["dofs"]) for b in self.boundaries]
        return np.unique(np.concatenate(d)) if d else np.zeros(0, dtype=np.int32)

    def unique_nodes(self, name: str) -> np.ndarray:
        return np.unique(np.concatenate([c.reshape(-1) for _, c in self.groups[name]]))


def _material_params(material: dict, model: int) -> np.ndarray:
    """bulk, shear, yield stress, hardening of one material (isotropic_elastic_property.cpp:10-43,
    isotropic_elastic_plastic.cpp:10-36)."""
    if "bulk_modulus" in material and "shear_modulus" in material:
        K, G = float(material["bulk_modulus"]), float(material["shear_modulus"])
    elif "elastic_modulus" in material and "poissons_ratio" in material:
        E, nu = float(material["elastic_modulus"]), float(material["poissons_ratio"])
        if nu >= 0.5:
            raise ValueError('"poissons_ratio" must be less than or equal to 0.5')
        K, G = E / (3.0 * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))
    else:
        raise ValueError('"elastic_modulus" and "poissons_ratio" or "bulk_modulus" and "shear_modulus" need to be '
                         "specified as material properties")
    sy = H = 0.0
    if model in (_capi.MODEL_J2, _capi.MODEL_J2_DAMAGE):
        if "yield_stress" not in material:
            raise ValueError('"yield_stress" was not specified in material properties')
        sy = float(material["yield_stress"])
        H = float(material.get("isotropic_hardening_modulus", 0.0))
    elif (model & 0xFF) >= _capi.MODEL_MICROSPHERE_GAUSSIAN:
        # micromechanical_elastomer.cpp:12-20, nonaffine_microsphere.cpp:14-24
        if "segments_per_chain" not in material:
            raise ValueError("segments_per_chain not specified in material data\n")
        sy = float(material["segments_per_chain"])
        if (model & 0xFF) == _capi.MODEL_MICROSPHERE_AGEING:
            # ageing_micromechanical_elastomer (micromechanical_elastomer.cpp:22-53): the
            # initial inactive shear modulus (1 - cure_time / 100) G
            H = (1.0 - _ageing_material(material)[2] / 100.0) * G
        if (model & 0xFF) == _capi.MODEL_MICROSPHERE_NONAFFINE:
            if "nonaffine_stretch_parameter" not in material:
                raise ValueError('"nonaffine_stretch_parameter" not specified in material data\n')
            H = float(material["nonaffine_stretch_parameter"])
    return np.array([K, G, sy, H])


DAMAGE_KEYS = ("kinematic_hardening_modulus", "softening_multiplier", "plasticity_viscous_denominator",
               "plasticity_viscous_exponent", "damage_viscous_denominator", "damage_viscous_exponent")


def _damage_material(material: dict) -> tuple:
    """The six constants of isotropic_elastic_plastic_damage (isotropic_elastic_plastic_damage.cpp:
    9-45) in neon_mesh_set_damage's order, with the reference's checks and messages (it checks
    them in the order softening, kinematic hardening, plasticity exponent and denominator, damage
    exponent and denominator)."""
    for key in ("softening_multiplier", "kinematic_hardening_modulus", "plasticity_viscous_exponent",
                "plasticity_viscous_denominator", "damage_viscous_exponent", "damage_viscous_denominator"):
        if key not in material:
            raise ValueError(f'"{key}" needs to be specified')
    return tuple(float(material[k]) for k in DAMAGE_KEYS)


def _ageing_material(material: dict) -> tuple:
    """(scission, recombination, cure_time) with the checks and messages of
    micromechanical_elastomer.cpp:22-53."""
    for key in ("scission_probability", "recombination_probability", "cure_time"):
        if key not in material:
            raise ValueError(f'"{key}" is not specified in "material" data')
    cure = float(material["cure_time"])
    if cure <= 10:
        raise ValueError('The "cure_time" is less than 10.  This is likely a mistake.  '
                         "Please provide a value between 10 and 100")
    ps, pr = float(material["scission_probability"]), float(material["recombination_probability"])
    if ps < 0.0 or pr < 0.0:
        raise ValueError("Material properties (probabilities) must be positive")
    return ps, pr, cure


def _microsphere(c: dict) -> int:
    """The microsphere branch of constitutive_model_factory.cpp:48-121, with its messages."""
    if "type" not in c:
        raise ValueError('Missing "type" as "affine" or "nonaffine" in microsphere model')
    if "quadrature" not in c:
        raise ValueError('Missing "quadrature" as "BO21", "BO33" or "FM900" in microsphere model')
    rules = {"BO21": 0, "BO33": _capi.SPHERE_BO33, "FM900": None}
    if c["quadrature"] not in rules:
        raise ValueError('"quadrature" field must be "BO21", "BO33" or "FM900"')
    rule = rules[c["quadrature"]]
    if c["type"] == "affine":
        if "statistics" not in c:
            raise ValueError('Missing "statistics" as "gaussian" or "langevin" in the microsphere model')
        if c["statistics"] not in ("gaussian", "langevin"):
            raise ValueError('"statistics" for the microsphere model must be either "gaussian" or "langevin"')
        if c["statistics"] == "gaussian" and "ageing" in c:
            if c["ageing"] != "BAND":
                raise ValueError('The only microsphere ageing model supported is "BAND"')
            if rule is None:
                raise ValueError("the FM900 unit-sphere rule does not run on the device path (BO21 and BO33 do)")
            return _capi.MODEL_MICROSPHERE_AGEING | rule
        model = _capi.MODEL_MICROSPHERE_GAUSSIAN if c["statistics"] == "gaussian" else _capi.MODEL_MICROSPHERE_LANGEVIN
    elif c["type"] == "nonaffine":
        model = _capi.MODEL_MICROSPHERE_NONAFFINE
    else:
        raise ValueError('microsphere model options are "affine" or "nonaffine"')
    if rule is None:
        raise ValueError("the FM900 unit-sphere rule does not run on the device path (BO21 and BO33 do)")
    return model | rule


def _model(constitutive: dict) -> int:
    """constitutive_model_factory.cpp:37-160, the models of the device path."""
    if "name" not in constitutive:
        raise ValueError('Missing "name" in "constitutive"')
    name = constitutive["name"]
    if name in ("neohooke", "compressible_neohooke"):
        return _capi.MODEL_NEOHOOKE
    if name == "microsphere":
        return _microsphere(constitutive)
    if name == "isotropic_linear_elasticity":
        return _capi.MODEL_LINEAR_ELASTIC
    if name == "J2_plasticity":
        if "finite_strain" not in constitutive:
            raise ValueError('"small_strain_J2_plasticity" must have a boolean value for "finite_strain"')
        if "damage" in constitutive:
            if constitutive["finite_strain"]:
                raise ValueError('"damage" is not yet implemented for "finite_strain"')
            if constitutive["damage"] == "isotropic_chaboche":
                return _capi.MODEL_J2_DAMAGE
            # any other damage type falls through to plain J2 (constitutive_model_factory.cpp:135-157)
        if constitutive["finite_strain"]:
            # finite_strain_J2_plasticity::consistent_tangent receives the logarithmic strain
            # where it expects the trial left Cauchy-Green tensor (finite_strain_J2_plasticity.cpp:
            # 101,139), so its tangent vanishes at the undeformed state
            raise ValueError("the finite-strain J2 model does not run on the device path")
        return _capi.MODEL_J2
    raise ValueError(f"The model name {name} is not recognised\n"
                     'Supported models are "neohooke", "microsphere" "isotropic_linear_elasticity" and '
                     '"J2_plasticity"\n')


def _read_mesh(path: str):
    with open(path) as fh:
        mesh = json.load(fh)
    if not mesh.get("Nodes"):
        raise ValueError('The mesh file is missing the "Nodes" field')
    if "Elements" not in mesh:
        raise ValueError('The mesh file is missing the "Elements" field')
    X = np.asarray(mesh["Nodes"][0]["Coordinates"], dtype=np.float64).reshape(-1, 3)
    groups: dict = {}
    for g in mesh["Elements"]:
        for key in ("Name", "Type", "NodalConnectivity"):
            if key not in g:
                raise ValueError(f'The element group in the mesh file is missing the "{key}" field')
        conn = np.asarray(g["NodalConnectivity"], dtype=np.int64)
        if conn.size == 0:
            raise ValueError("The element group in the mesh file is empty")
        if conn.min() < 0 or conn.max() >= len(X):
            raise ValueError(f'element group "{g["Name"]}" refers to nodes outside the mesh')
        groups.setdefault(g["Name"], []).append((int(g["Type"]), conn.astype(np.int32)))
    return X, groups


def _sinusoidal(boundary: dict, key: str, time_step: float):
    """boundary::generate_sinusoidal (src/mesh/boundary/boundary.cpp:87-143): from time 0
    (value 0), block i appends floor(cycles_i period_i / time_step) equal steps spanning
    cycles_i period_i, each valued amplitude_i sin(2 pi / period_i t + phase_i) at its absolute
    time t."""
    for field in ("period", "phase", "number_of_cycles"):
        if field not in boundary:
            raise ValueError(f'"{field}" was not specified in "boundaries".')
    amplitude, period = list(boundary[key]), list(boundary["period"])
    cycles, phase = list(boundary["number_of_cycles"]), list(boundary["phase"])
    if not (len(amplitude) == len(period) == len(phase) == len(cycles)):
        raise ValueError("The amplitude, period, number_of_cycles, time_step and phase vectors must be of the same "
                         "size.")
    times, loads = [0.0], [0.0]
    for a, T, c, ph in zip(amplitude, period, cycles, phase):
        omega = 2.0 * np.pi / T
        steps = int(c * T / time_step)  # the vector length truncates the quotient
        if steps < 1:
            raise ValueError(f'boundary "{boundary["name"]}": a sinusoidal block shorter than the time step')
        scale, start = c * T / steps, times[-1]
        block = [k * scale + start for k in range(1, steps + 1)]
        times += block
        loads += [a * np.sin(omega * t + ph) for t in block]
    return np.array(times), np.array(loads)


def _history(boundary: dict, key: str, time_step: float = 0.0):
    if "generate_type" in boundary:
        if boundary["generate_type"] != "sinusoidal":
            raise ValueError(f'"generate_type" {boundary["generate_type"]} is not recognised; use "sinusoidal"')
        return _sinusoidal(boundary, key, time_step)
    t = np.asarray(boundary["time"], dtype=np.float64)
    v = np.asarray(boundary[key], dtype=np.float64)
    if len(t) != len(v) or len(t) == 0:
        raise ValueError(f'boundary "{boundary["name"]}": "time" and "{key}" differ in length')
    return t, v


def _faces(groups: dict, name: str) -> np.ndarray:
    out = []
    for gtype, conn in groups[name]:
        if gtype != GMSH_QUAD4:
            raise ValueError(f'boundary "{name}": surface loads need quadrilateral4 faces (gmsh type 3), got {gtype}')
        out.append(conn)
    return np.concatenate(out)


def _sparse(vec: np.ndarray):
    nz = np.nonzero(vec)[0]
    return nz.astype(np.int32), vec[nz]


def read(path: str, step: int = 0) -> Problem:
    """Problem of load step `step` of the neon simulation file `path`."""
    from . import surface_load_vector, volume_load_vector

    with open(path) as fh:
        root = json.load(fh)
    for key in ("materials", "parts", "steps"):
        if key not in root:
            raise ValueError(f'"{key}" field does not exist in input file.')
    if len(root["parts"]) != 1:
        raise ValueError("the device path runs one part")
    part = root["parts"][0]
    materials = {m["name"]: m for m in root["materials"]}
    if part["material"] not in materials:
        raise ValueError(f'material "{part["material"]}" of part "{part["name"]}" is not defined')
    sim = root["steps"][step]
    for key in ("name", "time", "solution"):
        if key not in sim:
            raise ValueError(f'A simulation case needs a "{key}" field')
    if sim["solution"] not in ("equilibrium", "natural_frequency", "latin"):
        raise ValueError('"solution" is not valid.  Please use "equilibrium" or "natural_frequency"')
    if sim["solution"] == "latin":
        # module_factory.cpp:64-85 and latin_matrix.hpp:95-104
        if "nonlinear_options" not in sim:
            raise ValueError('"nonlinear_options" needs to be present for a solid_mechanics simulation')
        nl = sim["nonlinear_options"]
        if nl.get("latin_scheme") != "incremental" or nl.get("latin_local_stage") != "implicit":
            raise ValueError("Only incremental and implicit LATIN formulation has been implemented so far")
        if "latin_search_direction" not in nl:
            raise ValueError("latin_search_direction not specified in nonlinear_options")
    if sim["solution"] == "natural_frequency" and "eigen_solver" not in sim:
        raise ValueError('A natural frequency step needs an "eigen_solver" block')
    if len(sim["meshes"]) != 1:
        raise ValueError("Multibody simulation are not yet supported")
    mesh_data = sim["meshes"][0]
    if mesh_data["name"] != part["name"]:
        raise ValueError(f'Mesh name "{mesh_data["name"]}" was not found in the mesh store')
    model = _model(mesh_data.get("constitutive", {}))
    reduced = mesh_data.get("element_options", {}).get("quadrature", "full") == "reduced"
    if reduced:
        raise ValueError("the device path integrates hex8 elements with the full 8-point rule")

    X, groups = _read_mesh(os.path.join(os.path.dirname(os.path.abspath(path)), part["name"] + ".mesh"))
    if part["name"] not in groups:
        raise ValueError(f'Mesh name {part["name"]} does not exist in the mesh store')
    vols = groups[part["name"]]
    for gtype, _ in vols:
        if gtype != GMSH_HEX8:
            raise ValueError(f"the device path runs hexahedron8 elements (gmsh type 5), got type {gtype}")
    conn = np.concatenate([c for _, c in vols]).astype(np.int32)
    params = np.tile(_material_params(materials[part["material"]], model), (len(conn), 1))

    boundaries, loads = [], []
    # generated histories step with the first increment (solid_mechanics_module.cpp:15)
    dt0 = float(sim["time"]["increments"]["initial"]) if "increments" in sim.get("time", {}) else 0.0
    for b in mesh_data.get("boundaries", []):
        for key in ("name", "type"):
            if key not in b:
                raise ValueError(f'"{key}" was not specified in "boundary".')
        if "time" not in b and "generate_type" not in b:
            raise ValueError('Neither "time" nor "generate_type" was specified in "boundary".')
        name, btype = b["name"], b["type"]
        if name not in groups:
            raise ValueError(f"Mesh name {name} does not exist in the mesh store")
        if btype == "displacement":
            nodes = np.unique(np.concatenate([c.reshape(-1) for _, c in groups[name]]))
            for key, off in DOF_TABLE.items():
                if key in b:
                    t, v = _history(b, key, dt0)
                    boundaries.append({"name": name, "dofs": (3 * nodes + off).astype(np.int32), "times": t,
                                       "values": v})
        elif btype == "traction":
            faces = _faces(groups, name)
            for key, off in DOF_TABLE.items():
                if key in b:
                    t, v = _history(b, key, dt0)
                    d, vals = _sparse(surface_load_vector(X, faces, off, reduced))
                    loads.append({"name": name, "dofs": d, "values": vals, "times": t, "loads": v})
        elif btype == "pressure":
            if "value" not in b:
                raise ValueError('Pressure boundary condition must specify a "value" array')
            t, v = _history(b, "value", dt0)
            d, vals = _sparse(surface_load_vector(X, _faces(groups, name), -1, reduced))
            loads.append({"name": name, "dofs": d, "values": vals, "times": t, "loads": v})
        elif btype == "body_force":
            for gtype, _ in groups[name]:
                if gtype != GMSH_HEX8:
                    raise ValueError(f'boundary "{name}": body forces need hexahedron8 groups (gmsh type 5)')
            econn = np.concatenate([c for _, c in groups[name]])
            for key, off in DOF_TABLE.items():
                if key in b:
                    t, v = _history(b, key, dt0)
                    d, vals = _sparse(volume_load_vector(X, econn, off, reduced))
                    loads.append({"name": name, "dofs": d, "values": vals, "times": t, "loads": v})
        elif btype == "nodal_force":
            nodes = np.unique(np.concatenate([c.reshape(-1) for _, c in groups[name]]))
            for key, off in DOF_TABLE.items():
                if key in b:
                    t, v = _history(b, key, dt0)
                    loads.append({"name": name, "dofs": (3 * nodes + off).astype(np.int32),
                                  "values": np.ones(len(nodes)), "times": t, "loads": v, "fixed_time": 1.0})
        else:
            raise ValueError(f'boundary "{btype}" is not recognised')

    material = materials[part["material"]]
    density = np.full(len(conn), float(material["density"])) if "density" in material else None
    ageing = _ageing_material(material)[:2] if (model & 0xFF) == _capi.MODEL_MICROSPHERE_AGEING else None
    damage = _damage_material(material) if model == _capi.MODEL_J2_DAMAGE else None
    if sim["solution"] == "natural_frequency":
        if density is None:
            # material_property::initial_density (material_property.cpp:34-42)
            raise ValueError("density was requested, but not specified in the input file\n")
        return Problem(root.get("name", part["name"]), X, conn, params, model, boundaries, loads, {}, groups,
                       mesh_data.get("visualisation"), mesh_data["name"], sim["solution"], dict(sim["eigen_solver"]),
                       density, ageing, damage)
    ls = dict(sim.get("linear_solver", {}))
    if "type" not in ls:
        raise ValueError('A "linear_solver" "type" must be specified')
    if ls["type"] in DIRECT_SOLVERS:
        solver = {"tolerance": DIRECT_TOLERANCE, "maximum_iterations": DIRECT_MAX_ITERATIONS}
    elif ls["type"] == "iterative":
        solver = {"tolerance": ls.get("tolerance", 1.0e-5), "maximum_iterations": ls.get("maximum_iterations", 2000)}
    else:
        raise ValueError(f'Linear solver {ls["type"]} is not recognised.  Please use "PaStiX", "MUMPS", '
                         '"direct" or "iterative"')
    simulation = {"time": sim["time"], "nonlinear_options": sim.get("nonlinear_options", {}), "linear_solver": solver}
    return Problem(root.get("name", part["name"]), X, conn, params, model, boundaries, loads, simulation, groups,
                   mesh_data.get("visualisation"), mesh_data["name"], sim["solution"], None, density, ageing, damage)
