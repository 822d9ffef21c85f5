"""GPU world queries against the oracle's brute force, bit for bit: raycast hit lists
(bodies and entry distances, sorted by distance then body), closest hits, and AABB
queries, over the LBVH the last step built or one built on demand after an upload."""
import numpy as np
import pytest

from test_gpu_parity import _assert_same

pytestmark = pytest.mark.gpu


def _random_rays(rng, lo, hi, n, bodies_pos=None):
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    if bodies_pos is not None:  # half the rays aimed at a body centre
        k = n // 2
        tgt = bodies_pos[rng.integers(len(bodies_pos), size=k)]
        d[:k] = tgt - o[:k] + rng.normal(size=(k, 3)) * 0.05
    axis = rng.random(n) < 0.15  # axis-parallel rays (0 * inf in the slab test)
    d[axis] = 0.0
    d[axis, rng.integers(3, size=axis.sum())] = rng.choice([-1.0, 1.0], axis.sum())
    m = np.where(rng.random(n) < 0.5, np.inf, rng.uniform(0.5, (hi - lo), n))
    return o, d, m


def _check_rays(g, o_world, o, d, m, world=None, id_base=0, oracles=None):
    off, ids, dist = g.raycast(o, d, m, world)
    cid, cd = g.raycast_closest(o, d, m, world)
    hits = 0
    for r in range(len(o)):
        ow = o_world if oracles is None else oracles[world[r]]
        base = id_base if oracles is None else id_base[world[r]]
        oi, od = ow.raycast(o[r], d[r], m[r])
        gi = ids[off[r]:off[r + 1]].astype(np.int64) - base
        _assert_same(r, "ray hit ids", gi, oi.astype(np.int64))
        _assert_same(r, "ray hit distances", dist[off[r]:off[r + 1]], od)
        if len(oi):
            assert cid[r] == oi[0] + base and cd[r] == od[0], r
        else:
            assert cid[r] == 0xFFFFFFFF and cd[r] == np.inf, r
        hits += len(oi)
    return hits


def _check_boxes(g, o_world, boxes):
    off, ids = g.query_aabb(boxes)
    total = 0
    for q in range(len(boxes)):
        oi = o_world.query_aabb(boxes[q, :3], boxes[q, 3:])
        _assert_same(q, "aabb hits", ids[off[q]:off[q + 1]], oi)
        total += len(oi)
    return total


def test_raycast_and_aabb_query_boxes_scene(pk, oracle):
    s = pk.scene(0, None, 0)
    o_world = oracle.OracleWorld(iterations=s["iterations"], solver_mode=1, flags=s["flags"])
    o_world.load_scene(s)
    g = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"])
    g.load_scene(s)
    rng = np.random.default_rng(31)
    print("[queries] boxes scene: before any step (tree built on demand)", flush=True)
    o, d, m = _random_rays(rng, -12.0, 22.0, 1500, s["pos"])
    assert _check_rays(g, o_world, o, d, m) > 500
    for k in range(45):
        g.step(s["dt"])
        o_world.step(s["dt"])
    print("[queries] boxes scene: after 45 steps (fat bounds, the step's tree)", flush=True)
    b = g.bodies()
    _assert_same(45, "pos", b["pos"], o_world.bodies()["pos"])
    o, d, m = _random_rays(rng, -12.0, 22.0, 3000, b["pos"])
    assert _check_rays(g, o_world, o, d, m) > 1000
    # rays starting on a fat box face, parallel to it
    fat = b["fat"][rng.integers(len(b["fat"]), size=200)]
    o2 = np.stack([fat[:, 0] - 3.0, fat[:, 4], 0.5 * (fat[:, 2] + fat[:, 5])], axis=1)
    d2 = np.tile([1.0, 0.0, 0.0], (200, 1))
    _check_rays(g, o_world, o2, d2, np.full(200, np.inf))
    c = rng.uniform(-12, 22, (800, 3))
    h = rng.uniform(0.05, 4.0, (800, 3))
    boxes = np.concatenate([c - h, c + h], axis=1)
    assert _check_boxes(g, o_world, boxes) > 500
    # the Python world facade's single-ray form
    g.close()


def test_raycast_hull_clusters(pk, oracle):
    s = pk.scene(1, None, 20000)
    o_world = oracle.OracleWorld(iterations=s["iterations"], solver_mode=1, flags=s["flags"])
    o_world.load_scene(s)
    g = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"])
    g.load_scene(s)
    for _ in range(2):
        g.step(s["dt"])
        o_world.step(s["dt"])
    rng = np.random.default_rng(32)
    lo, hi = s["pos"].min(), s["pos"].max()
    o, d, m = _random_rays(rng, lo, hi, 600, s["pos"])
    print(f"[queries] hull clusters n={len(s['pos'])}", flush=True)
    assert _check_rays(g, o_world, o, d, m) > 300
    g.close()


def test_queries_batched_worlds(pk, oracle):
    """Per-world queries in a batched context: each query sees only its world's bodies."""
    import batched_util as bu
    s = pk.scene(pk.SCENE_BATCHED, 778, 5)
    fb = s["first_body"]
    g = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"])
    g.load_scene(s)
    g.set_worlds(fb)
    oracles = []
    for w in range(len(fb) - 1):
        ws, b0, b1 = bu.world_slice(s, w)
        o = oracle.OracleWorld(iterations=ws["iterations"], solver_mode=1, flags=ws["flags"])
        o.load_scene(ws)
        oracles.append(o)
    for _ in range(20):
        g.step(s["dt"])
        for o in oracles:
            o.step(s["dt"])
    rng = np.random.default_rng(33)
    pos = g.bodies()["pos"]
    n = 1200
    world = rng.integers(len(fb) - 1, size=n).astype(np.uint32)
    o_, d_, m_ = _random_rays(rng, pos.min(), pos.max(), n, pos)
    print(f"[queries] batched worlds sizes={np.diff(fb).tolist()}", flush=True)
    hits = _check_rays(g, None, o_, d_, m_, world=world, id_base=fb[:-1].astype(np.int64), oracles=oracles)
    assert hits > 200
    c = rng.uniform(pos.min(), pos.max(), (400, 3))
    h = rng.uniform(0.1, 3.0, (400, 3))
    boxes = np.concatenate([c - h, c + h], axis=1)
    qw = rng.integers(len(fb) - 1, size=400).astype(np.uint32)
    off, ids = g.query_aabb(boxes, qw)
    for q in range(400):
        oi = oracles[qw[q]].query_aabb(boxes[q, :3], boxes[q, 3:]).astype(np.int64) + fb[qw[q]]
        _assert_same(q, "batched aabb hits", ids[off[q]:off[q + 1]].astype(np.int64), oi)
    with pytest.raises(pk.PkgError):
        g.raycast([[0, 0, 0]], [[1, 0, 0]], np.inf, [len(fb) - 1])  # no such world
    g.close()


def test_queries_edge_cases(pk, oracle):
    """A single body, a disabled slot, a zero direction, an empty batch."""
    box = oracle.box_verts((0.5, 0.5, 0.5))

    def scene(n, static):
        return {"shape_offsets": np.array([0, 8], np.uint32), "verts": box,
                "is_static": np.array(static, np.uint8), "shape": np.zeros(n, np.uint32),
                "pos": np.array([[3.0 * i, 0, 0] for i in range(n)]), "vel": np.zeros((n, 3)),
                "angvel": np.zeros((n, 3)), "quat": np.tile([1.0, 0, 0, 0], (n, 1)), "mass": np.ones(n),
                "inertia": np.tile(np.eye(3).reshape(9), (n, 1)), "restitution": np.full(n, 0.5),
                "friction": np.full(n, 0.5)}

    for s in (scene(1, [1]), scene(3, [0, 2, 1])):
        g = pk.GpuWorld(gravity=(0, 0, 0))
        g.load_scene(s)
        o = oracle.OracleWorld(gravity=(0, 0, 0))
        o.load_scene(s)
        rays_o = np.array([[-5.0, 0, 0], [0, 0, 0], [3.0, 5.0, 0], [0.2, 0.1, 0.0]])
        rays_d = np.array([[1.0, 0, 0], [0, 0, 1.0], [0, -1.0, 0], [0.0, 0.0, 0.0]])
        _check_rays(g, o, rays_o, rays_d, np.full(4, np.inf))
        g.step(1 / 60)
        o.step(1 / 60)
        _check_rays(g, o, rays_o, rays_d, np.full(4, 2.0))
        _check_boxes(g, o, np.array([[-10.0, -1, -1, 10, 1, 1], [2.9, -1, -1, 3.1, 1, 1]]))
        off, ids, dist = g.raycast(np.zeros((0, 3)), np.zeros((0, 3)))
        assert len(off) == 1 and len(ids) == 0
        g.close()


def test_python_world_raycast(pk):
    w = pk.world(pk.world_desc.make().with_gravity((0, 0, 0)))
    box = pk.mesh.box((0.5, 0.5, 0.5))
    a = w.create_rigid(pk.object_desc.stat().with_mesh(box).with_pos((3, 0, 0)))
    b = w.create_rigid(pk.object_desc.stat().with_mesh(box).with_pos((7, 0, 0)))
    w.create_rigid(pk.object_desc.stat().with_mesh(box).with_pos((0, 10, 0)))
    assert w.raycast((0, 0, 0), (1, 0, 0), 100.0) == [(a, 2.5), (b, 6.5)]
    assert w.query_aabb((2, -1, -1), (8, 1, 1)) == [a, b]
    w.gpu.close()
