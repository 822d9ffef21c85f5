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
