"""Scene files: the JSON schema of PhysKit's demo and test helper
(tests/helper/src/graphics.cpp:20-135, the files under tests/configs/) read into the
structure-of-arrays scene that GpuWorld.load_scene uploads.

What the reference's reader does, and so what this one does:
  * world: gravity (default 0, -9.81, 0) and solver_iterations (default 10); camera,
    window and colour fields are display settings and are ignored;
  * objects: pos, vel (0), orientation [w, x, y, z] (identity), angular_velocity (0),
    inertia_tensor as the diagonal [Ixx, Iyy, Izz] (1, 1, 1), type "stat" | "dynam"
    (dynam), mass, restitution (0.5), friction (0.5), and a mesh:
      {"name": "box", "half_extents": [x, y, z]}           (default half extents 0.5)
      {"name": "pyramid", "height": h, "base_size": b}
      {"name": "sphere", "radius": r, "stacks": 16, "sectors": 32}
    Mesh sizes are read as 32-bit floats (graphics.cpp:24-41) and widened to double.
  * The reader passes (height, base_size) to mesh::pyramid(base_half, height)
    (graphics.cpp:104-106 against mesh.h:236), so a scene's "height" becomes the
    pyramid's base half-width and its "base_size" the apex height. This reader does
    the same, so a scene builds the same bodies as in the reference.
  * Identical mesh descriptions share one shape (graphics.cpp:91-110).
Numbers are parsed as doubles; mesh sizes are then rounded to float. The reference's
reader rounds the decimal text to float directly; the two agree unless a decimal lies
within half a double ulp of a float rounding boundary.
"""
from __future__ import annotations

import json
import math

import numpy as np


def _f32(x) -> float:
    return float(np.float32(x))


def _box(h):
    hx, hy, hz = h
    return [[-hx, -hy, -hz], [+hx, -hy, -hz], [+hx, +hy, -hz], [-hx, +hy, -hz],
            [-hx, -hy, +hz], [+hx, -hy, +hz], [+hx, +hy, +hz], [-hx, +hy, +hz]]


def _sphere(r, stacks, sectors):  # mesh::sphere, src/mesh.cpp:60-115
    v = [[0.0, r, 0.0]]
    for i in range(1, stacks):
        phi = math.pi * i / stacks
        for j in range(sectors):
            th = 2.0 * math.pi * j / sectors
            v.append([r * math.sin(phi) * math.cos(th), r * math.cos(phi), r * math.sin(phi) * math.sin(th)])
    v.append([0.0, -r, 0.0])
    return v


def _pyramid(base_half, height):  # mesh::pyramid, src/mesh.cpp:117-137
    b = base_half
    return [[-b, 0.0, -b], [b, 0.0, -b], [b, 0.0, b], [-b, 0.0, b], [0.0, height, 0.0]]


def mesh_vertices(m: dict) -> list:
    """Vertices of one mesh description, as the reference's reader builds them."""
    name = m.get("name", "box")
    if name == "box":
        return _box([_f32(x) for x in m.get("half_extents", (0.5, 0.5, 0.5))])
    if name == "pyramid":
        return _pyramid(_f32(m["height"]), _f32(m["base_size"]))
    if name == "sphere":
        stacks, sectors = int(m.get("stacks", 16)), int(m.get("sectors", 32))
        if stacks < 2 or sectors < 3:  # mesh.cpp:62-63
            raise ValueError("sphere needs stacks >= 2 and sectors >= 3")
        return _sphere(_f32(m["radius"]), stacks, sectors)
    raise ValueError(f"unknown mesh name {name!r}")


def parse(cfg: dict) -> dict:
    """The scene dict of a parsed scene file (see load)."""
    objs = cfg.get("objects", [])
    shapes, shape_of, verts, offs = {}, [], [], [0]
    for o in objs:
        m = o.get("mesh", {"name": "box", "half_extents": [0.5, 0.5, 0.5]})
        key = json.dumps(m, sort_keys=True)
        if key not in shapes:
            v = mesh_vertices(m)
            shapes[key] = len(offs) - 1
            verts.extend(v)
            offs.append(offs[-1] + len(v))
        shape_of.append(shapes[key])
    n = len(objs)
    for i, o in enumerate(objs):
        if "pos" not in o:
            raise ValueError(f"object {i}: pos is required")
    is_static = np.zeros(n, np.uint8)
    mass = np.zeros(n)
    for i, o in enumerate(objs):
        t = o.get("type", "dynam")
        if t not in ("stat", "dynam"):
            raise ValueError(f"object {i}: type must be 'stat' or 'dynam'")
        is_static[i] = 1 if t == "stat" else 0
        if "mass" not in o and t == "dynam":
            raise ValueError(f"object {i}: a dynamic object needs a mass")
        mass[i] = float(o.get("mass", 0.0))
    inertia = np.zeros((n, 9))
    for i, o in enumerate(objs):
        d = o.get("inertia_tensor", (1.0, 1.0, 1.0))
        inertia[i, [0, 4, 8]] = [float(x) for x in d]

    def arr(key, default, width):
        return np.array([[float(x) for x in o.get(key, default)] for o in objs], float).reshape(n, width)

    return {
        "shape_offsets": np.array(offs, np.uint32),
        "verts": np.array(verts, float).reshape(-1, 3),
        "is_static": is_static,
        "shape": np.array(shape_of, np.uint32),
        "pos": arr("pos", None, 3),
        "vel": arr("vel", (0.0, 0.0, 0.0), 3),
        "angvel": arr("angular_velocity", (0.0, 0.0, 0.0), 3),
        "quat": arr("orientation", (1.0, 0.0, 0.0, 0.0), 4),
        "mass": mass,
        "inertia": inertia,
        "restitution": np.array([float(o.get("restitution", 0.5)) for o in objs]),
        "friction": np.array([float(o.get("friction", 0.5)) for o in objs]),
        "n_static": int(is_static.sum()),
        "first_body": np.array([0, n], np.uint32),
        "gravity": tuple(float(x) for x in cfg.get("gravity", (0.0, -9.81, 0.0))),
        "iterations": int(cfg.get("solver_iterations", 10)),
        "title": cfg.get("title", ""),
    }


def load(path: str) -> dict:
    """Reads a scene file (tests/configs/*.json schema): the scene dict for
    GpuWorld.load_scene plus gravity, iterations and title."""
    with open(path) as f:
        return parse(json.load(f))
