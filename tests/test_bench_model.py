"""The bench's compulsory-byte models (CPU): closed forms on hand-built scenes, so a
dropped term or a miscounted pair shows up as a wrong number."""
import math
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
import bench  # noqa: E402


def _uv_sphere(r, stacks, sectors):
    v = [[0.0, r, 0.0]]
    for i in range(1, stacks):
        phi = math.pi * i / stacks
        for j in range(sectors):
            th = 2.0 * math.pi * j / sectors
            v.append([r * math.sin(phi) * math.cos(th), r * math.cos(phi), r * math.sin(phi) * math.sin(th)])
    v.append([0.0, -r, 0.0])
    return np.array(v)


def _box(h):
    return np.array([[sx * h, sy * h, sz * h] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])


def _scene(shapes, shape_of, pos):
    off = np.cumsum([0] + [len(v) for v in shapes]).astype(np.uint32)
    return {"shape_offsets": off, "verts": np.concatenate(shapes), "shape": np.array(shape_of, np.uint32),
            "pos": np.array(pos, float)}


def test_support_vertices_and_radii():
    sph = _uv_sphere(0.5, 16, 32)
    rnd = np.random.default_rng(5).normal(size=(100, 3))  # 100 vertices, not a UV sphere
    s = _scene([_box(0.5), sph, rnd], [0, 1, 2], [[0, 0, 0]] * 3)
    assert len(sph) == 482
    np.testing.assert_array_equal(bench.support_vertices(s), [8.0, 12.0, 100.0])
    r = bench.shape_radii(s)
    assert abs(r[0] - 0.5 * math.sqrt(3)) < 1e-15
    assert abs(r[1] - 0.5) < 1e-12
    assert r[2] == np.sqrt((rnd ** 2).sum(1)).max()


def test_gjk_stage_bytes_closed_form():
    """Two box pairs: one whose bounding spheres are apart (key and status 9 bytes, and
    32 bytes for each of its two bodies), one overlapping (key 8 + 2 x (24 + 32 + 12) +
    24 x (8 + 8) vertices + status 1 = 529 bytes), and a hit fraction adding
    (4 + 17 x 8) per pair. Culled pairs sharing a body read that body once."""
    b = _box(0.5)
    s = _scene([b], [0, 0, 0, 0, 0], [[0, 0, 0], [0.9, 0, 0], [10, 0, 0], [20, 0, 0], [30, 0, 0]])
    pairs = np.array([(0 << 32) | 1, (2 << 32) | 3], np.uint64)
    got = bench.algorithmic_bytes("gjk", s, pairs, 0.5)
    assert got == 9 + 2 * 32 + 529 + 0.5 * 2 * (4 + 17 * 8)
    shared = np.array([(2 << 32) | 3, (2 << 32) | 4], np.uint64)
    assert bench.algorithmic_bytes("gjk", s, shared, 0.0) == 2 * 9 + 3 * 32
    epa = bench.algorithmic_bytes("epa", s, pairs, 0.5)
    assert epa == 0.5 * 2 * (4 + 8 + 2 * (24 + 32 + 12) + 24 * 16 + 17 * 8 + 80 + 1)


def test_ray_bytes_closed_form():
    assert bench.ray_algorithmic_bytes(10, 4) == 680 + 96 + 228


def _counters(**kw):
    c = {k: 0 for k in ("moved", "candidates", "pairs", "gjk_hits", "contacts", "rows", "gjk_listed")}
    c.update(kw)
    return c


def test_stage_bytes_integrate_fusion_and_kept_list():
    """Four dynamic boxes and one static, 10 steps in one call: the first step runs
    k_integrate_vel (2n + 412 N), the other nine k_integrate_pos_vel (2n + 612 N), and
    only the last k_integrate_pos (n + 448 N); a step with no moved leaf and no
    candidate prices the pair update at n (the list is kept as it is)."""
    b = _box(0.5)
    s = _scene([b], [0] * 5, [[3.0 * i, 0, 0] for i in range(5)])
    s["n_static"] = 1
    n, nd = 5, 4
    c = _counters(pairs=3)
    pairs = np.zeros(0, np.uint64)
    out = bench.stage_bytes(s, c, pairs, 0.0, s["pos"], 10, False, steps=10)
    assert out["integrate_vel"][1] == ((2 * n + 412 * nd) + 9 * (2 * n + 612 * nd)) / 10 + 48 * 1
    assert out["integrate_pos"][1] == (n + 448 * nd) / 10
    assert out["pair_update"][1] == n
    one = bench.stage_bytes(s, c, pairs, 0.0, s["pos"], 10, False, steps=1)
    assert one["integrate_vel"][1] == 2 * n + 412 * nd + 48 and one["integrate_pos"][1] == n + 448 * nd
    moved = bench.stage_bytes(s, _counters(pairs=3, moved=2, candidates=1), pairs, 0.0, s["pos"], 10, False)
    assert moved["pair_update"][1] == 48 * 2 + 80 * n + 36 * 1 + 20 * 3
    off = bench.stage_bytes(s, c, pairs, 0.0, s["pos"], 10, False, integrate=False)
    assert off["integrate_vel"][1] == n and off["integrate_pos"][1] == 0


def test_stage_bytes_manifold_forms():
    """Below 2 pairs per body the manifold update runs its own pass over every pair
    (8 P) before the L listed pairs (28 L); from 2 pairs per body the cull settles the
    culled pairs, so the update prices only the listed pairs and the GJK stage carries
    6 P."""
    b = _box(0.5)
    s = _scene([b], [0] * 5, [[10.0 * i, 0, 0] for i in range(5)])
    s["n_static"] = 0
    few = np.array([(0 << 32) | 1], np.uint64)
    c = _counters(pairs=1, gjk_listed=1, contacts=2, gjk_hits=1)
    out = bench.stage_bytes(s, c, few, 1.0, s["pos"], 10, False)
    assert out["manifold"][1] == 8 * 1 + 28 * 1 + 208 * 2 + 192 * 1
    assert out["gjk"][1] == bench.algorithmic_bytes("gjk", s, few, 1.0, s["pos"])
    every = np.array([(a << 32) | c_ for a in range(5) for c_ in range(a + 1, 5)], np.uint64)  # 10 = 2 per body
    c10 = _counters(pairs=10, gjk_listed=3, contacts=2, gjk_hits=1)
    out10 = bench.stage_bytes(s, c10, every, 0.1, s["pos"], 10, False)
    assert out10["manifold"][1] == 28 * 3 + 208 * 2 + 192 * 1
    assert out10["gjk"][1] == bench.algorithmic_bytes("gjk", s, every, 0.1, s["pos"]) + 6 * 10
