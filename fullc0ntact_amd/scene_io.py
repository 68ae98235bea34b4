"""Scene files of the reference's drivers: the XML world parameters read by
FileParserXML::parseDataXML (inshape3dcore/preprocessing/reader.cpp:17-345,
WorldParameters in worldparameters.h:49-265) and the ASCII VTK particle
dump of CVtkWriter::WriteGPUParticleFile (vtkwriter.cpp:3399-3446).

Only what a particle-step driver consumes is mapped onto the simulation
parameters (gravity, time step, step count, domain extents, default
radius); every attribute the reference reads is parsed and kept, and the
rigid-body list is returned as plain records.
"""
from __future__ import annotations

import math
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import check, lib


def _atoi(s: str) -> int:
    """C atoi: the leading integer of s after whitespace, 0 if none."""
    s = s.strip()
    i = 1 if s[:1] in "+-" else 0
    j = i
    while j < len(s) and s[j].isdigit():
        j += 1
    return int(s[:j]) if j > i else 0


def _atof(s: str) -> float:
    """C atof: the longest leading decimal floating-point prefix, 0 if none."""
    s = s.strip()
    best = 0.0
    for end in range(len(s), 0, -1):
        try:
            best = float(s[:end])
            return best
        except ValueError:
            continue
    return best


def _floats(s: str, k: int) -> list[float]:
    """Stream extraction of k floats (missing values stay 0)."""
    out = []
    for tok in s.split()[:k]:
        try:
            out.append(float(tok))
        except ValueError:
            break
    return out + [0.0] * (k - len(out))


@dataclass
class RigidBodyRecord:
    """BodyStorage fields set from a <RigidBody> element."""
    shape_id: int = 0
    position: tuple = (0.0, 0.0, 0.0)
    velocity: tuple = (0.0, 0.0, 0.0)
    angular_velocity: tuple = (0.0, 0.0, 0.0)
    orientation: tuple = (0.0, 0.0, 0.0)
    force: tuple = (0.0, 0.0, 0.0)
    torque: tuple = (0.0, 0.0, 0.0)
    box_extents: tuple = (0.0, 0.0, 0.0)
    uvw: tuple = ((0.0, 0.0, 0.0),) * 3
    density: float = 0.0
    restitution: float = 0.0
    affected_by_gravity: int = 0
    mesh_file: str = ""
    dynamics_type: str = ""
    volume: float = 0.0
    tensor_diagonal: tuple = (0.0, 0.0, 0.0)


@dataclass
class WorldParameters:
    """WorldParameters with the constructor's defaults (worldparameters.cpp)."""
    start_type: int = 0
    bodies: int = 0
    body_init: int = 0
    total_timesteps: int = 0
    time_step: float = 0.01
    refinement_level: int = 0
    lcp_solver_iterations: int = 0
    pipeline_iterations: int = 0
    boundary_components: int = 0
    solver_type: int = 0
    liquid_solid: int = 0
    default_density: float = 1.5
    default_radius: float = 0.05
    liquid_density: float = 1.0
    air_friction: float = 1.0
    solution_file: str = "solution/particles.start0"
    body_file: str = "defaultBodies.i3d"
    ode_config_file: str = ""
    cgal_config_file: str = ""
    gravity: tuple = (0.0, 0.0, -10.0)
    extents: tuple = (-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)  # xmin xmax ymin ymax zmin zmax
    has_extents: bool = False
    do_fbm: bool = True
    do_dynamics: bool = True
    output_rigid_bodies: bool = False
    exclude_default_boundary: bool = False
    boundary_shapes: list = field(default_factory=list)  # (type, meshfile)
    rigid_bodies: list = field(default_factory=list)


_INT = {"starttype": "start_type", "liquidsolid": "liquid_solid", "nbodies": "bodies",
        "bodyinit": "body_init", "refinementlevel": "refinement_level",
        "totaltimesteps": "total_timesteps", "solvertype": "solver_type",
        "lcpsolveriterations": "lcp_solver_iterations",
        "collpipelineiterations": "pipeline_iterations"}
_FLOAT = {"defaultdensity": "default_density", "liquiddensity": "liquid_density",
          "airfriction": "air_friction", "defaultradius": "default_radius",
          "timestep": "time_step"}
_STR = {"solution": "solution_file", "bodyfile": "body_file", "odeconfigfile": "ode_config_file",
        "cgalconfigfile": "cgal_config_file"}


def read_world_xml(path) -> WorldParameters:
    """parseDataXML: ApplicationSettings attributes (names compared in lower
    case), the optional BoundaryDescription, and the RigidBodyList."""
    root = ET.parse(str(path)).getroot()
    w = WorldParameters()
    app = root.find("ApplicationSettings")
    if app is None:
        raise ValueError(f"{path}: no ApplicationSettings element")
    for name, value in app.attrib.items():
        key = name.lower()
        if key in _INT:
            setattr(w, _INT[key], _atoi(value))
        elif key in _FLOAT:
            setattr(w, _FLOAT[key], _atof(value))
        elif key in _STR:
            setattr(w, _STR[key], value)
        elif key == "gravity":
            w.gravity = tuple(_floats(value, 3))
        elif key == "dofbm":
            if _atoi(value) == 0:
                w.do_fbm = False
        elif key == "dodynamics":
            if _atoi(value) == 0:
                w.do_dynamics = False
        elif key == "outputrigidbodies":
            if _atoi(value) == 1:
                w.output_rigid_bodies = True
        elif key == "excludedefaultboundary":
            if _atoi(value) == 1:
                w.exclude_default_boundary = True
        elif key == "extents":
            w.extents = tuple(_floats(value, 6))
            w.has_extents = True
    bd = root.find("BoundaryDescription")
    if bd is not None:
        for name, value in bd.attrib.items():
            if name == "ncomponents":
                w.boundary_components = _atoi(value)
        for shape in bd.findall("BoundaryShape"):
            w.boundary_shapes.append((_atoi(shape.attrib.get("type", "0")),
                                      shape.attrib.get("meshfile", "")))
    rbl = root.find("RigidBodyList")
    if rbl is not None:
        for node in rbl:
            b = RigidBodyRecord()
            uvw = [(0.0, 0.0, 0.0)] * 3
            for name, value in node.attrib.items():
                key = name.lower()
                if key == "type":
                    b.shape_id = _atoi(value)
                elif key in ("position", "velocity", "angularvelocity", "orientation", "force",
                             "torque", "boxextents"):
                    attr = {"angularvelocity": "angular_velocity",
                            "boxextents": "box_extents"}.get(key, key)
                    setattr(b, attr, tuple(_floats(value, 3)))
                elif key in ("u", "v", "w"):
                    uvw["uvw".index(key)] = tuple(_floats(value, 3))
                elif key == "density":
                    b.density = _atof(value)
                elif key == "restitution":
                    b.restitution = _atof(value)
                elif key == "affectedbygravity":
                    b.affected_by_gravity = _atoi(value)
                elif key == "meshfile":
                    b.mesh_file = value
                elif key == "dynamicstype":
                    b.dynamics_type = value
                elif key == "volume":
                    b.volume = _atof(value)
                elif key == "tensor":
                    b.tensor_diagonal = tuple(_floats(value, 3))
            b.uvw = tuple(uvw)
            w.rigid_bodies.append(b)
    return w


def sim_params_from_world(world: WorldParameters, n: int, grid: int, radius=None,
                          **overrides) -> tuple["_lib.SimParams", float]:
    """Simulation parameters and time step of a particle step driven by a
    scene file: gravity, the particle radius (the scene's default radius
    unless given) with cell size 2r, and the domain's lower corner as the
    grid origin when the scene gives extents."""
    from .particle_system import make_params
    r = float(world.default_radius if radius is None else radius)
    P = make_params(n, grid, r, **overrides)
    P.gravity = type(P.gravity)(*[float(g) for g in world.gravity])
    if world.has_extents:
        e = world.extents
        P.worldOrigin = type(P.worldOrigin)(float(e[0]), float(e[2]), float(e[4]))
    return P, float(world.time_step)


@dataclass
class PluginRun:
    """What the FullC0ntact particle driver takes from a scene file
    (Particles_GPUTest, applications/particles_gpu/particles_gpu.cpp:
    44-218, 264-283): the number of pipeline steps, the VTK output cadence
    and the rigid bodies whose distance maps bound the particles.  The
    particle parameters are not in the scene: cuda_initParticles
    (cuda_addon/particledem_um.cu:280-330) fixes them, and the scene's
    gravity drives only the rigid-body world (particles_gpu.cpp:186)."""
    params: "_lib.SimParams"
    dt: float
    steps: int
    output_every: int
    rigid_bodies: list


def plugin_run_from_world(world: WorldParameters, n: int) -> PluginRun:
    """The particle step of the CollisionPipeline<dem_gpu_unified> driver for
    a scene: n spheres of radius 1/64 on a 64^3 grid from (-1, -1, -1),
    spring .5, damping .02, shear .1, no attraction, gravity -0.0003 in z,
    boundary damping -.5, global damping 1 and dt .5 (particledem_um.cu:
    285-297); the run loop steps while the time-step counter is at most
    totalTimesteps (particles_gpu.cpp:266) and writes the particle VTK dump
    every 250 steps (:278-279).  The plug-in's evalForces has no cursor
    sphere, so it is moved out of reach."""
    from .particle_system import make_params
    P = make_params(n, 64, 1.0 / 64.0, colliderPos=(1.0e4, 1.0e4, 1.0e4), colliderRadius=0.0)
    return PluginRun(P, 0.5, int(world.total_timesteps), 250, list(world.rigid_bodies))


def write_particle_vtk(path, pos4, radius: float = 1.0 / 64.0, radius_from_w: bool = False):
    """WriteGPUParticleFile for host float4 positions (n,4).  The reference
    writes radius 1/64 for every sphere (the default here)."""
    p = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
    check(lib.fc_write_particle_vtk(str(Path(path)).encode(), p.ctypes.data, len(p),
                                    float(radius), 1 if radius_from_w else 0))
