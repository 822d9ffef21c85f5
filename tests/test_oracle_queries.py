"""Pins of the oracle's world queries on the CPU: ray::intersect_distance
(include/physkit/collision/bvh.h:58-91), world_base::raycast (world.h:259-319) and the
tree's query_aabb (bvh.h:313-344).

Checks: the reference's own tree query and raycast cases (tests/dynamic_bvh/main.cpp:
211-455, re-expressed over world bodies), closed-form entry distances, the slab
convention for rays parallel to a face, and a geometric property for random rays:
the reported point lies on the box surface and the ray is outside just before it."""
import math

import numpy as np

RNG = np.random.default_rng(404)


def _world(oracle, centers, half=0.5, static=True):
    box = oracle.box_verts((half, half, half))
    n = len(centers)
    s = {"shape_offsets": np.array([0, 8], np.uint32), "verts": box,
         "is_static": np.full(n, 1 if static else 0, np.uint8), "shape": np.zeros(n, np.uint32),
         "pos": np.array(centers, float), "vel": np.zeros((n, 3)), "angvel": np.zeros((n, 3)),
         "quat": np.tile([1.0, 0, 0, 0], (n, 1)), "mass": np.ones(n), "inertia": np.tile(np.eye(3).reshape(9), (n, 1)),
         "restitution": np.full(n, 0.5), "friction": np.full(n, 0.5)}
    w = oracle.OracleWorld(gravity=(0, 0, 0))
    w.load_scene(s)
    return w, s


def test_ray_box_closed_forms(oracle):
    rb = oracle.ray_box
    assert rb((0, 0, 0), (1, 0, 0), (4.5, -0.5, -0.5), (5.5, 0.5, 0.5)) == 4.5
    assert rb((0, 0, 0), (1, 0, 0), (4.5, 9.5, -0.5), (5.5, 10.5, 0.5)) is None       # off axis
    assert rb((0, 0, 0), (1, 0, 0), (49.5, -0.5, -0.5), (50.5, 0.5, 0.5), 10.0) is None  # beyond max
    assert rb((0, 0, 0), (1, 0, 0), (4.5, -0.5, -0.5), (5.5, 0.5, 0.5), 4.5) == 4.5    # max inclusive
    assert rb((0, 0, 0), (-1, 0, 0), (-5.5, -0.5, -0.5), (-4.5, 0.5, 0.5)) == 4.5     # negative direction
    assert rb((0, 0, 0), (1, 0, 0), (-5.5, -0.5, -0.5), (-4.5, 0.5, 0.5)) is None     # behind the origin
    assert rb((0, 0, 0), (2, 0, 0), (-1, -1, -1), (1, 1, 1)) == 0.0                   # origin inside
    t = rb((0, 0, 0), (1, 1, 1), (4, 4, 4), (6, 6, 6))                                 # direction normalized
    assert abs(t - 4 * math.sqrt(3)) < 1e-12
    # a ray parallel to a face: origin on the face plane counts as inside that slab
    assert rb((0, 0.5, 0), (1, 0, 0), (4.5, -0.5, -0.5), (5.5, 0.5, 0.5)) == 4.5
    assert rb((0, 0.5 + 1e-9, 0), (1, 0, 0), (4.5, -0.5, -0.5), (5.5, 0.5, 0.5)) is None
    assert rb((0, -0.5, 0), (1, 0, -0.0), (4.5, -0.5, -0.5), (5.5, 0.5, 0.5)) == 4.5


def test_ray_entry_point_lies_on_the_surface(oracle):
    """Random rays and boxes: a hit at t > 0 puts o + t d on the box surface with the
    ray outside just before; a miss never passes through the box interior."""
    hits = 0
    for _ in range(3000):
        c = RNG.uniform(-5, 5, 3)
        h = RNG.uniform(0.1, 2.0, 3)
        mn, mx = c - h, c + h
        o = RNG.uniform(-8, 8, 3)
        d = RNG.normal(size=3) if RNG.random() < 0.3 else (c + RNG.uniform(-1.5, 1.5, 3) * h) - o
        if RNG.random() < 0.2:
            d[RNG.integers(3)] = 0.0
        if not np.any(d):
            continue
        u = d / np.linalg.norm(d)
        t = oracle.ray_box(o, d, mn, mx)
        inside = lambda p, eps: np.all(p >= mn - eps) and np.all(p <= mx + eps)
        if t is None:
            ts = np.linspace(0, 30, 3001)
            pts = o + ts[:, None] * u
            assert not np.any(np.all((pts > mn + 1e-9) & (pts < mx - 1e-9), axis=1))
            continue
        hits += 1
        p = o + t * u
        assert inside(p, 1e-9)
        if t > 0:
            on_face = np.isclose(p, mn, atol=1e-9) | np.isclose(p, mx, atol=1e-9)
            assert on_face.any()
            assert not inside(o + (t - 1e-6) * u, 1e-9)
        else:
            assert inside(o, 1e-12)
    assert hits > 1000


def test_world_raycast_reference_cases(oracle):
    """tests/dynamic_bvh/main.cpp:331-455 over world bodies (unit boxes)."""
    w, _ = _world(oracle, [[3, 0, 0], [7, 0, 0], [12, 0, 0], [0, 10, 0]])
    ids, dist = w.raycast((0, 0, 0), (1, 0, 0), 100.0)
    assert list(ids) == [0, 1, 2] and list(dist) == [2.5, 6.5, 11.5]
    assert len(w.raycast((0, 0, 0), (1, 0, 0), 2.0)[0]) == 0          # too short to reach
    assert list(w.raycast((0, 0, 0), (1, 0, 0), 2.5)[0]) == [0]       # reaches exactly
    assert list(w.raycast((0, 0, 0), (0, 1, 0))[0]) == [3]
    assert list(w.raycast((0, 0, 0), (-1, 0, 0))[0]) == []
    w2, _ = _world(oracle, [[5, 5, 5]], half=1.0)
    ids, dist = w2.raycast((0, 0, 0), (1, 1, 1))
    assert list(ids) == [0] and abs(dist[0] - 4 * math.sqrt(3)) < 1e-12
    # ties at distance 0 (origin inside several boxes) come in body order
    w3, _ = _world(oracle, [[0.2, 0, 0], [0, 0, 0], [-0.2, 0, 0]])
    ids, dist = w3.raycast((0, 0, 0), (0, 0, 1))
    assert list(ids) == [0, 1, 2] and not dist.any()


def test_world_query_aabb_reference_cases(oracle):
    """tests/dynamic_bvh/main.cpp:211-311 over world bodies."""
    w, _ = _world(oracle, [[0, 0, 0], [10, 10, 10], [0.5, 0.5, 0.5]])
    assert list(w.query_aabb((-0.75, -0.75, -0.75), (1.25, 1.25, 1.25))) == [0, 2]
    assert list(w.query_aabb((99.5, 99.5, 99.5), (100.5, 100.5, 100.5))) == []
    assert list(w.query_aabb((0.5, -1, -1), (2, 1, 1))) == [0, 2]  # touching faces intersect
    centers = RNG.uniform(-50, 50, (200, 3))
    w2, s = _world(oracle, centers)
    q = (np.full(3, -10.0), np.full(3, 10.0))
    brute = [i for i in range(200) if np.all(centers[i] - 0.5 <= q[1]) and np.all(centers[i] + 0.5 >= q[0])]
    assert list(w2.query_aabb(*q)) == brute


def test_queries_see_fat_bounds_and_skip_disabled_slots(oracle):
    """Dynamic leaves are their fat boxes (0.1 m margin plus the predicted motion,
    bvh.cpp:475-514) and removed slots are not in the tree."""
    box = oracle.box_verts((0.5, 0.5, 0.5))
    s = {"shape_offsets": np.array([0, 8], np.uint32), "verts": box,
         "is_static": np.array([0, 2, 1], np.uint8), "shape": np.zeros(3, np.uint32),
         "pos": np.array([[0.0, 0, 0], [5.0, 0, 0], [0, -5.0, 0]]), "vel": np.array([[1.0, 0, 0], [0, 0, 0], [0, 0, 0]]),
         "angvel": np.zeros((3, 3)), "quat": np.tile([1.0, 0, 0, 0], (3, 1)), "mass": np.ones(3),
         "inertia": np.tile(np.eye(3).reshape(9), (3, 1)), "restitution": np.full(3, 0.5),
         "friction": np.full(3, 0.5)}
    w = oracle.OracleWorld(gravity=(0, 0, 0), flags=2)
    w.load_scene(s)
    for _ in range(3):
        w.step(1 / 60)
    fat = w.bodies()["fat"][0]
    pos = w.bodies()["pos"][0]
    assert fat[4] > pos[1] + 0.5  # margin above the exact box
    y = pos[1] + 0.55             # misses the exact box, meets the fat one
    ids, dist = w.raycast((-10, y, 0), (1, 0, 0))
    assert list(ids) == [0] and dist[0] == fat[0] - (-10)
    assert list(w.raycast((5, 10, 0), (0, -1, 0))[0]) == []  # the disabled slot
    assert list(w.query_aabb((4, -1, -1), (6, 1, 1))) == []
    assert list(w.query_aabb((-1, -6, -1), (1, 1, 1))) == [0, 2]


def test_cpu_tree_closest_hit_equals_brute_force(oracle):
    """The CPU tree walk (the CPU baseline of the raycast bench) returns the first
    element of the brute-force hit list for every ray."""
    centers = RNG.uniform(-30, 30, (3000, 3))
    w, _ = _world(oracle, centers, half=0.7)
    o = RNG.uniform(-40, 40, (1500, 3))
    d = centers[RNG.integers(3000, size=1500)] - o
    d[::7] = RNG.normal(size=(len(d[::7]), 3))
    d[::11, 1] = 0.0
    m = np.where(RNG.random(1500) < 0.5, np.inf, RNG.uniform(1, 60, 1500))
    ids, dist, sec = w.raycast_closest_bvh(o, d, m)
    for r in range(1500):
        bi, bd = w.raycast(o[r], d[r], m[r])
        if len(bi):
            assert ids[r] == bi[0] and dist[r] == bd[0], r
        else:
            assert ids[r] == 0xFFFFFFFF and dist[r] == np.inf, r
