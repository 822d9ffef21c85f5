"""ctypes binding of the oracle (oracle/_build/liboracle.so), the FP64 CPU
restatement of PhysKit's world step. Test infrastructure only: used by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# PKO_LIB names another build of the oracle: the planted-mistake test
# (tests/test_oracle_mutants.py) runs the pin tests against altered copies
LIB_PATH = os.environ.get("PKO_LIB") or os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_lib = None

D = C.POINTER(C.c_double)
U8 = C.POINTER(C.c_uint8)
U32 = C.POINTER(C.c_uint32)
U64 = C.POINTER(C.c_uint64)
I32 = C.POINTER(C.c_int32)


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.pko_world_new.restype = C.c_void_p
        L.pko_world_new.argtypes = [D, C.c_int, C.c_int, C.c_int]
        L.pko_world_free.argtypes = [C.c_void_p]
        L.pko_add_shape.restype = C.c_int
        L.pko_add_shape.argtypes = [C.c_void_p, D, C.c_int]
        L.pko_add_bodies.restype = C.c_int
        L.pko_add_bodies.argtypes = [C.c_void_p, C.c_int, U8, U32, D, D, D, D, D, D, D, D]
        L.pko_step.restype = C.c_double
        L.pko_step.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.pko_num_bodies.restype = C.c_int
        L.pko_num_bodies.argtypes = [C.c_void_p]
        L.pko_get_bodies.argtypes = [C.c_void_p, D, D, D, D, D, U8]
        L.pko_get_derived.argtypes = [C.c_void_p, D, D, D]
        L.pko_num_pairs.restype = C.c_size_t
        L.pko_num_pairs.argtypes = [C.c_void_p]
        L.pko_get_pairs.argtypes = [C.c_void_p, U64]
        L.pko_get_manifolds.argtypes = [C.c_void_p, I32, D]
        L.pko_get_events.restype = C.c_size_t
        L.pko_get_events.argtypes = [C.c_void_p, U64, I32, C.c_size_t]
        L.pko_get_counters.argtypes = [C.c_void_p, U64]
        L.pko_get_colours.restype = C.c_size_t
        L.pko_get_colours.argtypes = [C.c_void_p, U32, U64, C.c_size_t]
        L.pko_gjk_epa.restype = C.c_int
        L.pko_gjk_epa.argtypes = [C.c_int, D, D, C.c_int, C.c_int, D, D, C.c_int, D, I32]
        L.pko_mesh_box.restype = C.c_int
        L.pko_mesh_box.argtypes = [D, D]
        L.pko_mesh_sphere.restype = C.c_int
        L.pko_mesh_sphere.argtypes = [C.c_double, C.c_int, C.c_int, D]
        L.pko_mesh_pyramid.restype = C.c_int
        L.pko_mesh_pyramid.argtypes = [C.c_double, C.c_double, D]
        L.pko_instance_bounds.argtypes = [D, C.c_int, D, D, D]
        L.pko_instance_support.argtypes = [D, C.c_int, D, D, D, D]
        L.pko_obb_support.argtypes = [D, D, D]
        L.pko_sincos.argtypes = [D, D, D, C.c_int]
        L.pko_exp_map.argtypes = [D, C.c_double, D]
        L.pko_inverse3.argtypes = [D, D]
        L.pko_set_bodies.argtypes = [C.c_void_p, C.c_int, U32, D, D, D, D]
        L.pko_raycast.restype = C.c_size_t
        L.pko_raycast.argtypes = [C.c_void_p, D, D, C.c_double, U32, D, C.c_size_t]
        L.pko_query_aabb.restype = C.c_size_t
        L.pko_query_aabb.argtypes = [C.c_void_p, D, U32, C.c_size_t]
        L.pko_raycast_closest_bvh.restype = C.c_double
        L.pko_raycast_closest_bvh.argtypes = [C.c_void_p, C.c_size_t, D, D, D, U32, D]
        L.pko_ray_box.restype = C.c_int
        L.pko_ray_box.argtypes = [D, D, D, C.c_double, D]
        L.pko_add_joint.restype = C.c_int
        L.pko_add_joint.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, D]
        L.pko_remove_joint.restype = C.c_int
        L.pko_remove_joint.argtypes = [C.c_void_p, C.c_uint32]
        L.pko_get_joints.restype = C.c_size_t
        L.pko_get_joints.argtypes = [C.c_void_p, U32, D, I32, C.c_size_t]
        L.pko_get_joint_colours.restype = C.c_size_t
        L.pko_get_joint_colours.argtypes = [C.c_void_p, U32, C.c_size_t]
        L.pko_run_stage.restype = C.c_int
        L.pko_run_stage.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.pko_set_manifold.restype = C.c_int
        L.pko_set_manifold.argtypes = [C.c_void_p, C.c_uint64, C.c_int, D]
        L.pko_set_fat.argtypes = [C.c_void_p, C.c_int, U32, D]
        L.pko_set_pairs.restype = C.c_int
        L.pko_set_pairs.argtypes = [C.c_void_p, C.c_size_t, U64, I32, D]
        L.pko_manifold_add.restype = C.c_int
        L.pko_manifold_add.argtypes = [C.c_int, D, D, D]
        _lib = L
    return _lib


def dp(a):
    return a.ctypes.data_as(D)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------- shapes for the standalone narrowphase ----------------
def aabb(mn, mx):
    return (0, f64(list(mn) + list(mx)), None)


def obb(center, quat, half):
    return (1, f64(list(center) + list(quat) + list(half)), None)


def mesh_instance(verts, pos=(0, 0, 0), quat=(1, 0, 0, 0)):
    return (2, f64(list(pos) + list(quat)), f64(verts).reshape(-1, 3))


def gjk_epa(a, b):
    """Returns None (separated), 'degenerate' (GJK hit, EPA padding failed), or a dict."""
    L = lib()
    out = np.zeros(10)
    st = np.zeros(4, dtype=np.int32)
    va = a[2] if a[2] is not None else np.zeros((1, 3))
    vb = b[2] if b[2] is not None else np.zeros((1, 3))
    r = L.pko_gjk_epa(a[0], dp(a[1]), dp(va), len(va), b[0], dp(b[1]), dp(vb), len(vb), dp(out),
                      st.ctypes.data_as(I32))
    if r == 0:
        return None
    if r == 2:
        return "degenerate"
    return {"normal": out[0:3].copy(), "world_a": out[3:6].copy(), "world_b": out[6:9].copy(),
            "depth": float(out[9]), "stats": st.copy()}


def box_verts(half):
    out = np.zeros((8, 3))
    lib().pko_mesh_box(dp(f64(half)), dp(out))
    return out


def sphere_verts(r, stacks=16, sectors=32):
    out = np.zeros((2 + (stacks - 1) * sectors, 3))
    n = lib().pko_mesh_sphere(r, stacks, sectors, dp(out))
    return out[:n]


def pyramid_verts(base_half, height):
    out = np.zeros((5, 3))
    lib().pko_mesh_pyramid(base_half, height, dp(out))
    return out


def instance_bounds(verts, pos, quat):
    out = np.zeros(6)
    v = f64(verts).reshape(-1, 3)
    lib().pko_instance_bounds(dp(v), len(v), dp(f64(pos)), dp(f64(quat)), dp(out))
    return out


def instance_support(verts, pos, quat, d):
    out = np.zeros(3)
    v = f64(verts).reshape(-1, 3)
    lib().pko_instance_support(dp(v), len(v), dp(f64(pos)), dp(f64(quat)), dp(f64(d)), dp(out))
    return out


def obb_support(center, quat, half, d):
    out = np.zeros(3)
    lib().pko_obb_support(dp(f64(list(center) + list(quat) + list(half))), dp(f64(d)), dp(out))
    return out


def sincos(x):
    x = f64(x)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    lib().pko_sincos(dp(x), dp(s), dp(c), len(x))
    return s, c


def exp_map(w, dt):
    q = np.zeros(4)
    lib().pko_exp_map(dp(f64(w)), dt, dp(q))
    return q


def ray_box(origin, direction, mn, mx, max_dist=float("inf")):
    """ray::intersect_distance (bvh.h:58-91): entry distance or None."""
    t = np.zeros(1)
    r = lib().pko_ray_box(dp(f64(origin)), dp(f64(direction)), dp(f64(list(mn) + list(mx))), float(max_dist), dp(t))
    return float(t[0]) if r else None


def manifold_add(old, new):
    """manifold::add_contact (collision_phases.h:124-130): old is (n, 13) contacts in the
    pko_get_manifolds layout, new one contact; returns the resulting (m, 13) contacts."""
    old = f64(np.asarray(old, float).reshape(-1, 13))
    out = np.zeros((4, 13))
    m = lib().pko_manifold_add(len(old), dp(old), dp(f64(np.asarray(new, float).reshape(13))), dp(out))
    assert m >= 0
    return out[:m].copy()


def inverse3(m):
    out = np.zeros(9)
    lib().pko_inverse3(dp(f64(m).reshape(9)), dp(out))
    return out.reshape(3, 3)


# ---------------- joints ----------------
JOINT_FIELDS = (("local_a", 3), ("local_b", 3), ("local_axis_a", 3), ("local_axis_b", 3),
                ("anchor", 3), ("axis", 3), ("distance", 1), ("stiffness", 1), ("damping", 1))


def joint_params(j):
    """The 21 doubles of pko_add_joint from a joint dict (missing fields are zero)."""
    out = []
    for name, n in JOINT_FIELDS:
        v = j.get(name, [0.0] * n if n > 1 else 0.0)
        out.extend(list(np.ravel(v)) if n > 1 else [float(v)])
    return f64(out)


# ---------------- worlds ----------------
class OracleWorld:
    """The oracle world. flags: 1 all leaves moved, 2 no solver, 4 no integration.
    solver_mode: 0 sequential in manifold order, 1 colour order (the device order)."""

    def __init__(self, gravity=(0.0, -9.81, 0.0), iterations=10, solver_mode=1, flags=0):
        self.L = lib()
        self.h = self.L.pko_world_new(dp(f64(gravity)), iterations, solver_mode, flags)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.pko_world_free(self.h)
            self.h = None

    def load_scene(self, scene):
        """scene: dict with shape_offsets (S+1,), verts (V,3) and per-body SoA arrays."""
        offs = scene["shape_offsets"]
        verts = f64(scene["verts"]).reshape(-1, 3)
        for s in range(len(offs) - 1):
            v = np.ascontiguousarray(verts[offs[s]:offs[s + 1]])
            self.L.pko_add_shape(self.h, dp(v), len(v))
        n = len(scene["shape"])
        is_static = np.ascontiguousarray(scene["is_static"], dtype=np.uint8)
        shape = np.ascontiguousarray(scene["shape"], dtype=np.uint32)
        r = self.L.pko_add_bodies(
            self.h, n, is_static.ctypes.data_as(U8), shape.ctypes.data_as(U32),
            dp(f64(scene["pos"])), dp(f64(scene["vel"])), dp(f64(scene["angvel"])),
            dp(f64(scene["quat"])), dp(f64(scene["mass"])), dp(f64(scene["inertia"])),
            dp(f64(scene["restitution"])), dp(f64(scene["friction"])))
        assert r == n, r

    def step(self, dt, nsteps=1):
        return self.L.pko_step(self.h, dt, nsteps)

    STAGES = {"integrate_vel": 0, "pairs": 1, "narrow": 2, "solve": 3, "integrate_pos": 4}

    def run_stage(self, name, dt=0.0):
        """One stage of world::step (src/world.cpp:20-56)."""
        assert self.L.pko_run_stage(self.h, self.STAGES[name], dt) == 1

    def set_state(self, bodies, pairs, counts, data):
        """Restores a saved state: every body's pos, vel, angvel, quat and fat (the dict of
        bodies()), the sorted pair keys and their manifolds (counts, (n, 4, 13) data)."""
        n = len(bodies["pos"])
        ids = np.arange(n, dtype=np.uint32)
        self.set_bodies(ids, pos=bodies["pos"], vel=bodies["vel"], angvel=bodies["angvel"], quat=bodies["quat"])
        fat = f64(bodies["fat"])
        self.L.pko_set_fat(self.h, n, ids.ctypes.data_as(U32), dp(fat))
        keys = np.ascontiguousarray(pairs, dtype=np.uint64)
        cnt = np.ascontiguousarray(counts, dtype=np.int32)
        d = f64(np.asarray(data, float).reshape(-1))
        assert self.L.pko_set_pairs(self.h, len(keys), keys.ctypes.data_as(U64), cnt.ctypes.data_as(I32),
                                    dp(d) if len(d) else None) == 1

    def set_manifold(self, key, contacts):
        """Replaces the cached manifold of active pair `key` by contacts ((n, 13), the
        layout of manifolds())."""
        c = f64(np.asarray(contacts, float).reshape(-1, 13))
        assert self.L.pko_set_manifold(self.h, int(key), len(c), dp(c) if len(c) else None) == 1

    def bodies(self):
        n = self.L.pko_num_bodies(self.h)
        pos, vel, ang = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        quat, fat = np.zeros((n, 4)), np.zeros((n, 6))
        moved = np.zeros(n, dtype=np.uint8)
        self.L.pko_get_bodies(self.h, dp(pos), dp(vel), dp(ang), dp(quat), dp(fat),
                              moved.ctypes.data_as(U8))
        return {"pos": pos, "vel": vel, "angvel": ang, "quat": quat, "fat": fat, "moved": moved}

    def derived(self):
        n = self.L.pko_num_bodies(self.h)
        iw, iwi, im = np.zeros((n, 9)), np.zeros((n, 9)), np.zeros(n)
        self.L.pko_get_derived(self.h, dp(iw), dp(iwi), dp(im))
        return iw, iwi, im

    def pairs(self):
        n = self.L.pko_num_pairs(self.h)
        k = np.zeros(n, dtype=np.uint64)
        if n:
            self.L.pko_get_pairs(self.h, k.ctypes.data_as(U64))
        return k

    def manifolds(self):
        n = self.L.pko_num_pairs(self.h)
        cnt = np.zeros(n, dtype=np.int32)
        data = np.zeros((n, 4, 13))
        if n:
            self.L.pko_get_manifolds(self.h, cnt.ctypes.data_as(I32), dp(data))
        return cnt, data

    def events(self):
        n = self.L.pko_get_events(self.h, None, None, 0)
        k = np.zeros(n, dtype=np.uint64)
        t = np.zeros(n, dtype=np.int32)
        if n:
            self.L.pko_get_events(self.h, k.ctypes.data_as(U64), t.ctypes.data_as(I32), n)
        return k, t

    def counters(self):
        out = np.zeros(6, dtype=np.uint64)
        self.L.pko_get_counters(self.h, out.ctypes.data_as(U64))
        return dict(zip(["moved", "pairs", "gjk_hits", "contacts", "rows", "colours"],
                        [int(x) for x in out]))

    def set_bodies(self, ids, pos=None, vel=None, angvel=None, quat=None):
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        arrs = [None if a is None else f64(a) for a in (pos, vel, angvel, quat)]
        self.L.pko_set_bodies(self.h, len(ids), ids.ctypes.data_as(U32),
                              *[None if a is None else dp(a) for a in arrs])

    def raycast(self, origin, direction, max_dist=float("inf")):
        """[(body, entry distance)] sorted by (distance, body)."""
        o, d = f64(origin), f64(direction)
        n = self.L.pko_raycast(self.h, dp(o), dp(d), float(max_dist), None, None, 0)
        ids = np.zeros(n, np.uint32)
        dist = np.zeros(n)
        if n:
            self.L.pko_raycast(self.h, dp(o), dp(d), float(max_dist), ids.ctypes.data_as(U32), dp(dist), n)
        return ids, dist

    def raycast_closest_bvh(self, origins, dirs, max_dist):
        """Closest hit per ray through a CPU tree walk: (ids, dists, seconds walking)."""
        o = f64(np.asarray(origins, float).reshape(-1, 3))
        d = f64(np.asarray(dirs, float).reshape(-1, 3))
        m = f64(np.broadcast_to(np.asarray(max_dist, float), (len(o),)))
        ids = np.zeros(len(o), np.uint32)
        dist = np.zeros(len(o))
        sec = self.L.pko_raycast_closest_bvh(self.h, len(o), dp(o), dp(d), dp(m), ids.ctypes.data_as(U32), dp(dist))
        return ids, dist, sec

    def query_aabb(self, mn, mx):
        b = f64(list(mn) + list(mx))
        n = self.L.pko_query_aabb(self.h, dp(b), None, 0)
        ids = np.zeros(n, np.uint32)
        if n:
            self.L.pko_query_aabb(self.h, dp(b), ids.ctypes.data_as(U32), n)
        return ids

    def add_joint(self, j):
        """j: a joint dict (type, slot, a, b and the desc fields; see joint_params)."""
        return bool(self.L.pko_add_joint(self.h, int(j["type"]), int(j.get("slot", 0)), int(j["a"]),
                                         int(j["b"]), dp(joint_params(j))))

    def remove_joint(self, key):
        return bool(self.L.pko_remove_joint(self.h, int(key)))

    def joints(self):
        n = self.L.pko_get_joints(self.h, None, None, None, 0)
        keys = np.zeros(n, dtype=np.uint32)
        acc = np.zeros((n, 6))
        rows = np.zeros(n, dtype=np.int32)
        if n:
            self.L.pko_get_joints(self.h, keys.ctypes.data_as(U32), dp(acc), rows.ctypes.data_as(I32), n)
        return keys, acc, rows

    def joint_colours(self):
        n = self.L.pko_get_joint_colours(self.h, None, 0)
        out = np.zeros(n, dtype=np.uint32)
        if n:
            self.L.pko_get_joint_colours(self.h, out.ctypes.data_as(U32), n)
        return out

    def colours(self, with_keys=False):
        n = self.L.pko_get_colours(self.h, None, None, 0)
        out = np.zeros(n, dtype=np.uint32)
        keys = np.zeros(n, dtype=np.uint64)
        if n:
            self.L.pko_get_colours(self.h, out.ctypes.data_as(U32), keys.ctypes.data_as(U64), n)
        return (out, keys) if with_keys else out
