"""physkit_amd: PhysKit's collision-and-contact world step on MI355X (gfx950).

Python mirror of the C ABI in include/physkit_amd.h (ctypes). It follows PhysKit's
world API (include/physkit/core/world.h:23-49, 202-242; object.h:15-109; mesh.h):
``world_desc``, ``object_desc.dynam()/stat()`` with the ``with_*`` builders,
``mesh.box/sphere/pyramid/make``, ``world.create_rigid`` and ``world.step(dt)``.
Every step runs on the GPU through libphyskit_amd.so; there is no CPU path. Loading
fails loudly when the library is missing or no GPU is present.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libphyskit_amd.so")

PKG_OK = 0
STATUS = {0: "PKG_OK", 1: "PKG_INVALID_ARG", 2: "PKG_OOM", 3: "PKG_OVERFLOW_PAIRS",
          4: "PKG_OVERFLOW_EPA", 5: "PKG_DEVICE_ERROR", 6: "PKG_NO_DEVICE"}
FLAG_ALL_MOVED, FLAG_NO_SOLVER, FLAG_NO_INTEGRATE, FLAG_EPA_GUARD, FLAG_FP32_SOLVE = 1, 2, 4, 8, 16
FLAG_FP32_GJK, FLAG_FP32_BROAD = 32, 64
# the mixed-precision form: FP32 broadphase leaf tests, certified FP32 GJK separation,
# FP32 contact solve (EPA, manifolds and body state stay FP64)
FLAG_MIXED = FLAG_FP32_SOLVE | FLAG_FP32_GJK | FLAG_FP32_BROAD

D = C.POINTER(C.c_double)
U8 = C.POINTER(C.c_uint8)
U32 = C.POINTER(C.c_uint32)
U64 = C.POINTER(C.c_uint64)
I32 = C.POINTER(C.c_int32)

EXPORTS = [
    "pkg_create", "pkg_destroy", "pkg_last_error", "pkg_upload_shapes", "pkg_upload_bodies",
    "pkg_set_body_state", "pkg_step", "pkg_step_timed", "pkg_download_bodies", "pkg_get_pairs",
    "pkg_get_manifolds", "pkg_get_events", "pkg_get_counters", "pkg_get_colours",
    "pkg_gjk_epa_batch", "pkg_scene_create", "pkg_scene_info_get", "pkg_scene_copy",
    "pkg_scene_destroy", "pkg_set_worlds", "pkg_scene_copy_worlds", "pkg_add_joints",
    "pkg_remove_joints", "pkg_get_joints", "pkg_raycast", "pkg_raycast_closest", "pkg_raycast_closest_device",
    "pkg_query_aabb", "pkg_gjk_epa_batch_device", "pkg_autotune", "pkg_get_tuning", "pkg_set_tuning",
]

# joint types, numbered as constraint_type (include/physkit/collision/constraint.h:14-23)
JOINT_DISTANCE, JOINT_BALL_SOCKET, JOINT_HINGE, JOINT_SLIDER, JOINT_WELD, JOINT_SPRING = range(6)
JOINT_DIMENSION = {JOINT_DISTANCE: 1, JOINT_BALL_SOCKET: 3, JOINT_HINGE: 5, JOINT_SLIDER: 5,
                   JOINT_WELD: 6, JOINT_SPRING: 1}


def joint_key(jtype: int, slot: int) -> int:
    """(type << 24) | slot: the solve order of joints (type, then arena slot)."""
    return (int(jtype) << 24) | int(slot)


class PkgError(RuntimeError):
    def __init__(self, status, msg=""):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class pkg_config(C.Structure):
    _fields_ = [("gravity", C.c_double * 3), ("solver_iterations", C.c_uint32), ("flags", C.c_uint32)]


class pkg_body_desc(C.Structure):
    _fields_ = [("is_static", U8), ("shape", U32), ("pos", D), ("vel", D), ("angvel", D),
                ("quat", D), ("mass", D), ("inertia", D), ("restitution", D), ("friction", D)]


class pkg_body_state(C.Structure):
    _fields_ = [("pos", D), ("vel", D), ("angvel", D), ("quat", D), ("fat", D), ("moved", U8)]


class pkg_counters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("moved", "pairs", "gjk_hits", "contacts", "rows",
                                          "colours", "epa_overflow", "candidates", "gjk_listed",
                                          "gjk_fp64", "steps_total", "pairs_total", "gjk_hits_total")]


class pkg_tuning(C.Structure):
    _fields_ = [("gjk_lanes", C.c_int32), ("epa_block", C.c_int32), ("epa_lds_max", C.c_uint32),
                ("tree_refits", C.c_uint32),
                ("gjk_ms", C.c_float * 4), ("epa_block_ms", C.c_float * 3), ("epa_lds_ms", C.c_float * 3)]


class pkg_joint_desc(C.Structure):
    _fields_ = [("type", C.c_int32), ("slot", C.c_uint32), ("a", C.c_uint32), ("b", C.c_uint32),
                ("local_a", C.c_double * 3), ("local_b", C.c_double * 3), ("local_axis_a", C.c_double * 3),
                ("local_axis_b", C.c_double * 3), ("anchor", C.c_double * 3), ("axis", C.c_double * 3),
                ("distance", C.c_double), ("stiffness", C.c_double), ("damping", C.c_double)]


def _joint_struct(j: dict) -> pkg_joint_desc:
    d = pkg_joint_desc()
    d.type, d.slot, d.a, d.b = int(j["type"]), int(j.get("slot", 0)), int(j["a"]), int(j["b"])
    for f in ("local_a", "local_b", "local_axis_a", "local_axis_b", "anchor", "axis"):
        v = np.asarray(j.get(f, (0.0, 0.0, 0.0)), dtype=float).reshape(3)
        getattr(d, f)[:] = [float(x) for x in v]
    d.distance = float(j.get("distance", 0.0))
    d.stiffness = float(j.get("stiffness", 0.0))
    d.damping = float(j.get("damping", 0.0))
    return d


class pkg_scene_info(C.Structure):
    _fields_ = [("n_shapes", C.c_uint32), ("n_verts", C.c_uint32), ("n_bodies", C.c_uint32),
                ("n_static", C.c_uint32), ("dt", C.c_double), ("iterations", C.c_uint32),
                ("steps", C.c_uint32), ("flags", C.c_uint32), ("n_worlds", C.c_uint32)]


_lib = None


def lib():
    """Loads libphyskit_amd.so (built by __graft_entry__.build / physkit_amd/build.py)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"physkit_amd native library not built: {LIB_PATH}")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    L.pkg_create.argtypes = [C.POINTER(P), C.POINTER(pkg_config)]
    L.pkg_destroy.argtypes = [P]
    L.pkg_last_error.argtypes = [P]
    L.pkg_last_error.restype = C.c_char_p
    L.pkg_upload_shapes.argtypes = [P, C.c_uint32, U32, D]
    L.pkg_upload_bodies.argtypes = [P, C.c_uint32, C.POINTER(pkg_body_desc)]
    L.pkg_set_body_state.argtypes = [P, C.c_uint32, U32, D, D, D, D]
    L.pkg_step.argtypes = [P, C.c_double, C.c_int]
    L.pkg_step_timed.argtypes = [P, C.c_double, C.c_int, D, C.c_int]
    L.pkg_autotune.argtypes = [P, C.c_double, C.c_int, C.POINTER(pkg_tuning)]
    L.pkg_get_tuning.argtypes = [P, C.POINTER(pkg_tuning)]
    L.pkg_set_tuning.argtypes = [P, C.POINTER(pkg_tuning)]
    L.pkg_download_bodies.argtypes = [P, C.POINTER(pkg_body_state)]
    L.pkg_get_pairs.argtypes = [P, U64, C.c_size_t, C.POINTER(C.c_size_t)]
    L.pkg_get_manifolds.argtypes = [P, I32, D, C.c_size_t, C.POINTER(C.c_size_t)]
    L.pkg_get_events.argtypes = [P, U64, I32, C.c_size_t, C.POINTER(C.c_size_t)]
    L.pkg_get_counters.argtypes = [P, C.POINTER(pkg_counters)]
    L.pkg_get_colours.argtypes = [P, U32, U64, C.c_size_t, C.POINTER(C.c_size_t)]
    L.pkg_set_worlds.argtypes = [P, C.c_uint32, U32]
    L.pkg_scene_copy_worlds.argtypes = [P, U32]
    L.pkg_gjk_epa_batch.argtypes = [P, C.c_uint32, I32, D, U32, U32, I32, D, U32, U32, D,
                                    C.c_uint32, D, I32]
    L.pkg_scene_create.argtypes = [C.POINTER(P), C.c_int, C.c_uint64, C.c_uint32]
    L.pkg_scene_info_get.argtypes = [P, C.POINTER(pkg_scene_info)]
    L.pkg_scene_copy.argtypes = [P, U32, D, U8, U32, D, D, D, D, D, D, D, D]
    L.pkg_scene_destroy.argtypes = [P]
    L.pkg_raycast.argtypes = [P, C.c_uint32, D, D, D, U32, U32, U32, D, C.c_size_t, C.POINTER(C.c_size_t)]
    L.pkg_raycast_closest.argtypes = [P, C.c_uint32, D, D, D, U32, U32, D]
    L.pkg_raycast_closest_device.argtypes = [P, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_void_p, D]
    L.pkg_query_aabb.argtypes = [P, C.c_uint32, D, U32, U32, U32, C.c_size_t, C.POINTER(C.c_size_t)]
    L.pkg_add_joints.argtypes = [P, C.POINTER(pkg_joint_desc), C.c_size_t]
    L.pkg_remove_joints.argtypes = [P, U32, C.c_size_t]
    L.pkg_get_joints.argtypes = [P, U32, D, C.c_size_t, C.POINTER(C.c_size_t)]
    for name in EXPORTS:
        if name != "pkg_last_error":
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(D)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ------------------------------------------------------------------ scenes
SCENE_BOXES, SCENE_HULLS, SCENE_PYRAMIDS, SCENE_BATCHED, SCENE_GRANULAR, SCENE_GRANULAR_LOW = 0, 1, 2, 3, 4, 5
SEED_BASE = 20261202


def scene(kind: int, seed: int | None = None, n_override: int = 0) -> dict:
    """Synthetic scene from the library's generators as numpy arrays (host input)."""
    L = lib()
    seed = SEED_BASE + kind if seed is None else seed
    h = C.c_void_p()
    st = L.pkg_scene_create(C.byref(h), kind, seed, n_override)
    if st != PKG_OK:
        raise PkgError(st, "scene")
    try:
        info = pkg_scene_info()
        L.pkg_scene_info_get(h, C.byref(info))
        nb, ns, nv = info.n_bodies, info.n_shapes, info.n_verts
        s = {
            "shape_offsets": np.zeros(ns + 1, np.uint32), "verts": np.zeros((nv, 3)),
            "is_static": np.zeros(nb, np.uint8), "shape": np.zeros(nb, np.uint32),
            "pos": np.zeros((nb, 3)), "vel": np.zeros((nb, 3)), "angvel": np.zeros((nb, 3)),
            "quat": np.zeros((nb, 4)), "mass": np.zeros(nb), "inertia": np.zeros((nb, 9)),
            "restitution": np.zeros(nb), "friction": np.zeros(nb),
        }
        L.pkg_scene_copy(h, s["shape_offsets"].ctypes.data_as(U32), _dp(s["verts"]),
                         s["is_static"].ctypes.data_as(U8), s["shape"].ctypes.data_as(U32),
                         _dp(s["pos"]), _dp(s["vel"]), _dp(s["angvel"]), _dp(s["quat"]),
                         _dp(s["mass"]), _dp(s["inertia"]), _dp(s["restitution"]),
                         _dp(s["friction"]))
        fb = np.zeros(info.n_worlds + 1, np.uint32)
        L.pkg_scene_copy_worlds(h, fb.ctypes.data_as(U32))
        s.update(dt=info.dt, iterations=info.iterations, steps=info.steps, flags=info.flags,
                 n_static=info.n_static, first_body=fb)
        return s
    finally:
        L.pkg_scene_destroy(h)


def scene_world(s: dict, w: int):
    """World w of a batched scene as a scene of its own (bodies [first_body[w],
    first_body[w+1]) renumbered from 0, with its contiguous shapes), and the body range
    it came from."""
    fb = s["first_body"]
    b0, b1 = int(fb[w]), int(fb[w + 1])
    shp = s["shape"][b0:b1].astype(np.int64)
    s0, s1 = int(shp.min()), int(shp.max()) + 1
    offs = s["shape_offsets"].astype(np.int64)
    out = {k: s[k][b0:b1].copy() for k in ("is_static", "pos", "vel", "angvel", "quat", "mass", "inertia",
                                            "restitution", "friction")}
    out["shape"] = (shp - s0).astype(np.uint32)
    out["shape_offsets"] = (offs[s0:s1 + 1] - offs[s0]).astype(np.uint32)
    out["verts"] = s["verts"][offs[s0]:offs[s1]].copy()
    for k in ("dt", "iterations", "steps", "flags"):
        out[k] = s[k]
    out["n_static"] = int(np.sum(out["is_static"] != 0))
    out["first_body"] = np.array([0, b1 - b0], np.uint32)
    return out, b0, b1


# ------------------------------------------------------------------ bulk world (SoA)
class GpuWorld:
    """One world on the GPU, loaded from SoA arrays (the bulk form of the C ABI)."""

    def __init__(self, gravity=(0.0, -9.81, 0.0), iterations=10, flags=0):
        self.L = lib()
        cfg = pkg_config()
        for k in range(3):
            cfg.gravity[k] = gravity[k]
        cfg.solver_iterations = iterations
        cfg.flags = flags
        h = C.c_void_p()
        st = self.L.pkg_create(C.byref(h), C.byref(cfg))
        if st != PKG_OK:
            raise PkgError(st, "pkg_create")
        self.h = h
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.pkg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, st, what):
        if st != PKG_OK:
            msg = self.L.pkg_last_error(self.h)
            raise PkgError(st, f"{what}: {msg.decode() if msg else ''}")

    def load_scene(self, s: dict):
        offs = np.ascontiguousarray(s["shape_offsets"], dtype=np.uint32)
        verts = _f64(s["verts"]).reshape(-1, 3)
        self._check(self.L.pkg_upload_shapes(self.h, len(offs) - 1, offs.ctypes.data_as(U32),
                                             _dp(verts)), "upload_shapes")
        n = len(s["shape"])
        keep = {k: _f64(s[k]) for k in ("pos", "vel", "angvel", "quat", "mass", "inertia",
                                        "restitution", "friction")}
        keep["is_static"] = np.ascontiguousarray(s["is_static"], dtype=np.uint8)
        keep["shape"] = np.ascontiguousarray(s["shape"], dtype=np.uint32)
        d = pkg_body_desc(keep["is_static"].ctypes.data_as(U8), keep["shape"].ctypes.data_as(U32),
                          _dp(keep["pos"]), _dp(keep["vel"]), _dp(keep["angvel"]),
                          _dp(keep["quat"]), _dp(keep["mass"]), _dp(keep["inertia"]),
                          _dp(keep["restitution"]), _dp(keep["friction"]))
        self._check(self.L.pkg_upload_bodies(self.h, n, C.byref(d)), "upload_bodies")
        self.n = n

    def step(self, dt, nsteps=1):
        self._check(self.L.pkg_step(self.h, dt, nsteps), "step")

    STAGES = ("integrate_vel", "lbvh_build", "pair_update", "gjk", "epa", "manifold",
              "solver", "integrate_pos", "solve_iterations")

    def step_timed(self, dt, nsteps=1):
        """Steps and returns the summed device milliseconds per stage (STAGES;
        solve_iterations is the part of solver spent in the iteration kernels)."""
        ms = np.zeros(len(self.STAGES))
        self._check(self.L.pkg_step_timed(self.h, dt, nsteps, _dp(ms), len(self.STAGES)), "step_timed")
        return dict(zip(self.STAGES, ms.tolist()))

    def autotune(self, dt, steps=1):
        """pkg_autotune: steps the world `steps` times, timing every launch-shape
        candidate of the GJK and EPA stages on each step's own inputs, and keeps the
        fastest; returns the choice and the measured stage milliseconds."""
        t = pkg_tuning()
        self._check(self.L.pkg_autotune(self.h, dt, steps, C.byref(t)), "autotune")
        return {"gjk_lanes": t.gjk_lanes, "epa_block": t.epa_block, "epa_lds_max": t.epa_lds_max,
                "tree_refits": t.tree_refits, "gjk_ms": [round(x, 4) for x in t.gjk_ms], "epa_block_ms": [round(x, 4) for x in t.epa_block_ms],
                "epa_lds_ms": [round(x, 4) for x in t.epa_lds_ms]}

    def tuning(self):
        t = pkg_tuning()
        self._check(self.L.pkg_get_tuning(self.h, C.byref(t)), "get_tuning")
        return {"gjk_lanes": t.gjk_lanes, "epa_block": t.epa_block, "epa_lds_max": t.epa_lds_max,
                "tree_refits": t.tree_refits}

    def set_tuning(self, gjk_lanes, epa_block, epa_lds_max, tree_refits=None):
        """pkg_set_tuning; tree_refits None keeps the current value."""
        if tree_refits is None:
            tree_refits = self.tuning()["tree_refits"]
        t = pkg_tuning(gjk_lanes=gjk_lanes, epa_block=epa_block, epa_lds_max=epa_lds_max, tree_refits=tree_refits)
        self._check(self.L.pkg_set_tuning(self.h, C.byref(t)), "set_tuning")

    def set_state(self, ids, pos=None, vel=None, angvel=None, quat=None):
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        arrs = [None if a is None else _f64(a) for a in (pos, vel, angvel, quat)]
        ptrs = [None if a is None else _dp(a) for a in arrs]
        self._check(self.L.pkg_set_body_state(self.h, len(ids), ids.ctypes.data_as(U32), *ptrs),
                    "set_body_state")

    def bodies(self):
        n = self.n
        out = {"pos": np.zeros((n, 3)), "vel": np.zeros((n, 3)), "angvel": np.zeros((n, 3)),
               "quat": np.zeros((n, 4)), "fat": np.zeros((n, 6)), "moved": np.zeros(n, np.uint8)}
        st = pkg_body_state(_dp(out["pos"]), _dp(out["vel"]), _dp(out["angvel"]),
                            _dp(out["quat"]), _dp(out["fat"]), out["moved"].ctypes.data_as(U8))
        self._check(self.L.pkg_download_bodies(self.h, C.byref(st)), "download_bodies")
        return out

    def pairs(self):
        n = C.c_size_t()
        self._check(self.L.pkg_get_pairs(self.h, None, 0, C.byref(n)), "get_pairs")
        k = np.zeros(n.value, np.uint64)
        if n.value:
            self._check(self.L.pkg_get_pairs(self.h, k.ctypes.data_as(U64), n.value, C.byref(n)),
                        "get_pairs")
        return k

    def manifolds(self):
        n = C.c_size_t()
        self._check(self.L.pkg_get_manifolds(self.h, None, None, 0, C.byref(n)), "get_manifolds")
        cnt = np.zeros(n.value, np.int32)
        data = np.zeros((n.value, 4, 13))
        if n.value:
            self._check(self.L.pkg_get_manifolds(self.h, cnt.ctypes.data_as(I32), _dp(data),
                                                 n.value, C.byref(n)), "get_manifolds")
        return cnt, data

    def events(self):
        n = C.c_size_t()
        self._check(self.L.pkg_get_events(self.h, None, None, 0, C.byref(n)), "get_events")
        k = np.zeros(n.value, np.uint64)
        t = np.zeros(n.value, np.int32)
        if n.value:
            self._check(self.L.pkg_get_events(self.h, k.ctypes.data_as(U64), t.ctypes.data_as(I32),
                                              n.value, C.byref(n)), "get_events")
        return k, t

    def colours(self, with_keys=False):
        n = C.c_size_t()
        self._check(self.L.pkg_get_colours(self.h, None, None, 0, C.byref(n)), "get_colours")
        out = np.zeros(n.value, np.uint32)
        keys = np.zeros(n.value, np.uint64)
        if n.value:
            self._check(self.L.pkg_get_colours(self.h, out.ctypes.data_as(U32), keys.ctypes.data_as(U64),
                                               n.value, C.byref(n)), "get_colours")
        return (out, keys) if with_keys else out

    def set_worlds(self, first_body):
        """Batched independent worlds: bodies [first_body[w], first_body[w+1]) form world w."""
        fb = np.ascontiguousarray(first_body, dtype=np.uint32)
        self._check(self.L.pkg_set_worlds(self.h, len(fb) - 1, fb.ctypes.data_as(U32)), "set_worlds")

    @staticmethod
    def _rays(origins, dirs, max_dist):
        o = _f64(np.asarray(origins, dtype=float).reshape(-1, 3))
        d = _f64(np.asarray(dirs, dtype=float).reshape(-1, 3))
        n = len(o)
        m = _f64(np.broadcast_to(np.asarray(max_dist, dtype=float), (n,)))
        return n, o, d, m

    @staticmethod
    def _worlds(world, n):
        if world is None:
            return None, None
        w = np.ascontiguousarray(np.broadcast_to(np.asarray(world, dtype=np.uint32), (n,)))
        return w, w.ctypes.data_as(U32)

    def raycast(self, origins, dirs, max_dist=np.inf, world=None):
        """world_base::raycast for a batch of rays: (offsets[n+1], ids, dists), the hits
        of ray i at [offsets[i], offsets[i+1]) sorted by (distance, body)."""
        n, o, d, m = self._rays(origins, dirs, max_dist)
        w, wp = self._worlds(world, n)
        off = np.zeros(n + 1, np.uint32)
        tot = C.c_size_t()
        self._check(self.L.pkg_raycast(self.h, n, _dp(o), _dp(d), _dp(m), wp, off.ctypes.data_as(U32), None, None,
                                       0, C.byref(tot)), "raycast")
        ids = np.zeros(tot.value, np.uint32)
        dist = np.zeros(tot.value)
        if tot.value:
            self._check(self.L.pkg_raycast(self.h, n, _dp(o), _dp(d), _dp(m), wp, off.ctypes.data_as(U32),
                                           ids.ctypes.data_as(U32), _dp(dist), tot.value, C.byref(tot)), "raycast")
        return off, ids, dist

    def raycast_closest(self, origins, dirs, max_dist=np.inf, world=None):
        """First hit per ray in (distance, body) order: (ids, dists); 0xFFFFFFFF / inf on a miss."""
        n, o, d, m = self._rays(origins, dirs, max_dist)
        w, wp = self._worlds(world, n)
        ids = np.zeros(n, np.uint32)
        dist = np.zeros(n)
        self._check(self.L.pkg_raycast_closest(self.h, n, _dp(o), _dp(d), _dp(m), wp, ids.ctypes.data_as(U32),
                                               _dp(dist)), "raycast_closest")
        return ids, dist

    def query_aabb(self, boxes, world=None):
        """query_aabb for a batch of boxes (min xyz, max xyz): (offsets[n+1], ids)."""
        b = _f64(np.asarray(boxes, dtype=float).reshape(-1, 6))
        n = len(b)
        w, wp = self._worlds(world, n)
        off = np.zeros(n + 1, np.uint32)
        tot = C.c_size_t()
        self._check(self.L.pkg_query_aabb(self.h, n, _dp(b), wp, off.ctypes.data_as(U32), None, 0, C.byref(tot)),
                    "query_aabb")
        ids = np.zeros(tot.value, np.uint32)
        if tot.value:
            self._check(self.L.pkg_query_aabb(self.h, n, _dp(b), wp, off.ctypes.data_as(U32),
                                              ids.ctypes.data_as(U32), tot.value, C.byref(tot)), "query_aabb")
        return off, ids

    def add_joints(self, joints):
        """joints: dicts with type, slot, a, b and the desc fields (pkg_joint_desc)."""
        arr = (pkg_joint_desc * max(len(joints), 1))(*[_joint_struct(j) for j in joints])
        self._check(self.L.pkg_add_joints(self.h, arr, len(joints)), "add_joints")

    def remove_joints(self, keys):
        k = np.ascontiguousarray(keys, dtype=np.uint32)
        self._check(self.L.pkg_remove_joints(self.h, k.ctypes.data_as(U32), len(k)), "remove_joints")

    def joints(self):
        """(keys, impulses[n, 6]) of the live joints in key order."""
        n = C.c_size_t()
        self._check(self.L.pkg_get_joints(self.h, None, None, 0, C.byref(n)), "get_joints")
        keys = np.zeros(n.value, np.uint32)
        imp = np.zeros((n.value, 6))
        if n.value:
            self._check(self.L.pkg_get_joints(self.h, keys.ctypes.data_as(U32), _dp(imp), n.value, C.byref(n)),
                        "get_joints")
        return keys, imp

    def counters(self):
        c = pkg_counters()
        self._check(self.L.pkg_get_counters(self.h, C.byref(c)), "get_counters")
        return {f: int(getattr(c, f)) for f, _ in pkg_counters._fields_}

    def gjk_epa_batch(self, a_list, b_list):
        """Standalone narrowphase over shape pairs. Shapes are tuples
        (kind, params10, verts or None): kind 0 aabb (min3,max3), 1 obb (center3,
        quat wxyz, half3), 2 mesh instance (pos3, quat wxyz) with its vertices."""
        n = len(a_list)
        verts = [np.zeros((0, 3))]
        off = 0
        cols = {}
        for side, lst in (("a", a_list), ("b", b_list)):
            kind = np.zeros(n, np.int32)
            prm = np.zeros((n, 10))
            vb = np.zeros(n, np.uint32)
            ve = np.zeros(n, np.uint32)
            for i, (k, p, v) in enumerate(lst):
                kind[i] = k
                p = np.asarray(p, dtype=np.float64).ravel()
                prm[i, :len(p)] = p
                if k == 2:
                    v = _f64(v).reshape(-1, 3)
                    vb[i] = off
                    off += len(v)
                    ve[i] = off
                    verts.append(v)
            cols[side] = (kind, prm, vb, ve)
        V = _f64(np.concatenate(verts) if off else np.zeros((1, 3)))
        out = np.zeros((n, 10))
        hit = np.zeros(n, np.int32)
        ka, pa, ba, ea = cols["a"]
        kb, pb, bb, eb = cols["b"]
        self._check(self.L.pkg_gjk_epa_batch(
            self.h, n, ka.ctypes.data_as(I32), _dp(pa), ba.ctypes.data_as(U32),
            ea.ctypes.data_as(U32), kb.ctypes.data_as(I32), _dp(pb), bb.ctypes.data_as(U32),
            eb.ctypes.data_as(U32), _dp(V), max(off, 1), _dp(out), hit.ctypes.data_as(I32)),
            "gjk_epa_batch")
        return hit, out


# ------------------------------------------------------------------ physkit-style API
def _box_verts(h):
    hx, hy, hz = h
    return np.array([[-hx, -hy, -hz], [+hx, -hy, -hz], [+hx, +hy, -hz], [-hx, +hy, -hz],
                     [-hx, -hy, +hz], [+hx, -hy, +hz], [+hx, +hy, +hz], [-hx, +hy, +hz]])


class mesh:
    """Convex mesh (src/mesh.cpp:16-137). Only the vertices enter the world step."""

    def __init__(self, vertices):
        self.vertices = _f64(vertices).reshape(-1, 3)
        if len(self.vertices) == 0:
            raise ValueError("mesh needs at least one vertex")

    @staticmethod
    def make(vertices, triangles=None):
        return mesh(vertices)

    @staticmethod
    def box(half_extents):
        return mesh(_box_verts(half_extents))

    @staticmethod
    def sphere(radius, stacks=16, sectors=32):
        v = [[0.0, radius, 0.0]]
        for i in range(1, stacks):
            phi = math.pi * i / stacks
            for j in range(sectors):
                th = 2.0 * math.pi * j / sectors
                v.append([radius * math.sin(phi) * math.cos(th), radius * math.cos(phi),
                          radius * math.sin(phi) * math.sin(th)])
        v.append([0.0, -radius, 0.0])
        return mesh(v)

    @staticmethod
    def pyramid(base_half, height):
        b = base_half
        return mesh([[-b, 0.0, -b], [b, 0.0, -b], [b, 0.0, b], [-b, 0.0, b], [0.0, height, 0.0]])


class world_desc:
    """include/physkit/core/world.h:23-49."""

    def __init__(self):
        self.gravity = (0.0, -9.81, 0.0)
        self.solver_iterations = 10
        self.flags = 0

    @staticmethod
    def make():
        return world_desc()

    def with_epa_guard(self, on=True):
        """Opt-in FLAG_EPA_GUARD (not in the reference): drop degenerate EPA results."""
        self.flags = (self.flags | FLAG_EPA_GUARD) if on else (self.flags & ~FLAG_EPA_GUARD)
        return self

    def with_gravity(self, g):
        self.gravity = tuple(float(x) for x in g)
        return self

    def with_solver_iterations(self, n):
        self.solver_iterations = int(n)
        return self


class object_desc:
    """include/physkit/core/object.h:15-109 (defaults: mass 1, identity inertia,
    restitution 0.5, friction 0.5)."""

    def __init__(self, static):
        self.static = static
        self.pos = (0.0, 0.0, 0.0)
        self.vel = (0.0, 0.0, 0.0)
        self.mass = 1.0
        self.orientation = (1.0, 0.0, 0.0, 0.0)
        self.ang_vel = (0.0, 0.0, 0.0)
        self.inertia = np.eye(3)
        self.mesh = None
        self.restitution = 0.5
        self.friction = 0.5

    @staticmethod
    def dynam():
        return object_desc(False)

    @staticmethod
    def stat():
        return object_desc(True)

    def _set(self, k, v):
        setattr(self, k, v)
        return self

    def with_pos(self, p): return self._set("pos", tuple(p))
    def with_vel(self, v): return self._set("vel", tuple(v))
    def with_mass(self, m): return self._set("mass", float(m))
    def with_orientation(self, q): return self._set("orientation", tuple(q))
    def with_ang_vel(self, w): return self._set("ang_vel", tuple(w))
    def with_inertia_tensor(self, I): return self._set("inertia", np.asarray(I, dtype=float))
    def with_mesh(self, m): return self._set("mesh", m)
    def with_restitution(self, e): return self._set("restitution", float(e))
    def with_friction(self, f): return self._set("friction", float(f))


class _joint_desc:
    """The desc structs of the six joint types (constraint.h:144-175, 239-285, 357-374,
    446-497, 585-609, 737-755): make(a, b) then with_* setters."""
    jtype = -1

    def __init__(self, a, b):
        self.a, self.b = int(a), int(b)
        self.fields = {}

    @classmethod
    def make(cls, a, b):
        return cls(a, b)

    def _set(self, k, v):
        self.fields[k] = v
        return self

    def as_dict(self, slot):
        return {"type": self.jtype, "slot": slot, "a": self.a, "b": self.b, **self.fields}


class distance_desc(_joint_desc):
    jtype = JOINT_DISTANCE
    def with_local_anchor_a(self, p): return self._set("local_a", tuple(p))
    def with_local_anchor_b(self, p): return self._set("local_b", tuple(p))
    def with_distance(self, d): return self._set("distance", float(d))


class spring_desc(distance_desc):
    jtype = JOINT_SPRING
    def with_stiffness(self, k): return self._set("stiffness", float(k))
    def with_damping(self, c): return self._set("damping", float(c))


class ball_socket_desc(_joint_desc):
    jtype = JOINT_BALL_SOCKET
    def with_anchor(self, p): return self._set("anchor", tuple(p))


class hinge_desc(_joint_desc):
    jtype = JOINT_HINGE
    def with_local_anchor_a(self, p): return self._set("local_a", tuple(p))
    def with_local_anchor_b(self, p): return self._set("local_b", tuple(p))
    def with_local_axis_a(self, v): return self._set("local_axis_a", tuple(v))
    def with_local_axis_b(self, v): return self._set("local_axis_b", tuple(v))


class slider_desc(_joint_desc):
    jtype = JOINT_SLIDER
    def with_anchor(self, p): return self._set("anchor", tuple(p))
    def with_axis(self, v): return self._set("axis", tuple(v))


class weld_desc(_joint_desc):
    jtype = JOINT_WELD
    def with_anchor(self, p): return self._set("anchor", tuple(p))


class world:
    """physkit::world on the GPU: bodies are collected by create_rigid and uploaded
    on the first step; handles are arena slot indices (body ids). Joints added with
    add_constraint are built at the next step from the poses at that time, as the
    reference runs add_constraint commands before step_impl (world.h:234-242)."""

    def __init__(self, desc: world_desc | None = None):
        self.desc = desc or world_desc()
        self._descs = []
        self._gpu = None
        self._dirty = True
        self.on_collision = None       # callback(a, b) at collision begin
        self.on_collision_exit = None  # callback(a, b) at collision end
        # per-type joint arenas (arena.h:64-80): generation per slot, LIFO free slots
        self._jgen = {t: [] for t in range(6)}
        self._jfree = {t: [] for t in range(6)}
        self._jlive = set()
        self._jadd = []     # (key, dict) waiting for the next step
        self._jdel = []     # keys waiting for the next step

    def add_constraint(self, d: _joint_desc):
        """world::add_constraint (world.h:387-395): returns (type, slot, generation)."""
        t = d.jtype
        if self._jfree[t]:
            slot = self._jfree[t].pop()
        else:
            slot = len(self._jgen[t])
            self._jgen[t].append(0)
        key = joint_key(t, slot)
        self._jlive.add(key)
        self._jadd.append((key, d.as_dict(slot)))
        return (t, slot, self._jgen[t][slot])

    def remove_constraint(self, h) -> bool:
        """world::remove_constraint (world.h:397-398); False for a stale handle."""
        t, slot, gen = h
        key = joint_key(t, slot)
        if key not in self._jlive or self._jgen[t][slot] != gen:
            return False
        self._jlive.discard(key)
        self._jgen[t][slot] += 1
        self._jfree[t].append(slot)
        pending = [x for x in self._jadd if x[0] == key]
        if pending:
            self._jadd.remove(pending[0])
        else:
            self._jdel.append(key)
        return True

    def _flush_joints(self):
        if self._jdel:
            self._gpu.remove_joints(self._jdel)
            self._jdel = []
        if self._jadd:
            self._gpu.add_joints([j for _, j in self._jadd])
            self._jadd = []

    def create_rigid(self, d: object_desc) -> int:
        if d.mesh is None:
            raise ValueError("Object must have a mesh")  # object.h:121
        self._descs.append(d)
        self._dirty = True
        return len(self._descs) - 1

    def _upload(self):
        # bodies already on the device keep the state they reached; new ones start
        # from their object_desc
        kept = self._gpu.bodies() if self._gpu is not None and self._gpu.n > 0 else None
        offs = [0]
        verts = []
        for d in self._descs:
            verts.append(d.mesh.vertices)
            offs.append(offs[-1] + len(d.mesh.vertices))
        n = len(self._descs)
        s = {
            "shape_offsets": np.array(offs, np.uint32), "verts": np.concatenate(verts),
            "is_static": np.array([d.static for d in self._descs], np.uint8),
            "shape": np.arange(n, dtype=np.uint32),
            "pos": np.array([d.pos for d in self._descs], float),
            "vel": np.array([d.vel for d in self._descs], float),
            "angvel": np.array([d.ang_vel for d in self._descs], float),
            "quat": np.array([d.orientation for d in self._descs], float),
            "mass": np.array([d.mass for d in self._descs], float),
            "inertia": np.array([np.asarray(d.inertia, float).reshape(9) for d in self._descs]),
            "restitution": np.array([d.restitution for d in self._descs], float),
            "friction": np.array([d.friction for d in self._descs], float),
        }
        if kept is not None:
            m = len(kept["pos"])
            for k in ("pos", "vel", "angvel", "quat"):
                s[k][:m] = kept[k]
        if self._gpu is None:
            self._gpu = GpuWorld(self.desc.gravity, self.desc.solver_iterations, self.desc.flags)
        self._gpu.load_scene(s)
        self._dirty = False

    def step(self, dt):
        if self._dirty:
            self._upload()
        self._flush_joints()
        self._gpu.step(dt)
        if self.on_collision or self.on_collision_exit:
            keys, kinds = self._gpu.events()
            for k, t in zip(keys.tolist(), kinds.tolist()):
                a, b = k >> 32, k & 0xFFFFFFFF
                if t == 1 and self.on_collision:
                    self.on_collision(a, b)
                elif t == 2 and self.on_collision_exit:
                    self.on_collision_exit(a, b)

    def state(self):
        if self._dirty:
            self._upload()
        return self._gpu.bodies()

    def raycast(self, origin, direction, max_dist=np.inf):
        """world_base::raycast (world.h:259-319): [(handle, distance)] sorted by distance."""
        if self._dirty:
            self._upload()
        off, ids, dist = self._gpu.raycast([origin], [direction], max_dist)
        return list(zip(ids.tolist(), dist.tolist()))

    def query_aabb(self, mn, mx):
        """Handles whose tree bounds intersect the box, ascending."""
        if self._dirty:
            self._upload()
        off, ids = self._gpu.query_aabb([list(mn) + list(mx)])
        return ids.tolist()

    @property
    def gpu(self) -> GpuWorld:
        if self._dirty:
            self._upload()
        return self._gpu
