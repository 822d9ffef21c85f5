"""Edge inputs of the world step, bit for bit against the oracle: a world with no
bodies, batched partitions with empty and one-body worlds, bodies so fast that pairs
appear and vanish every step and leaves travel far between full LBVH builds, and
drifting hull clusters whose contacts break while the pair list is kept whole."""
import numpy as np
import pytest

import batched_util as bu
from test_gpu_parity import _assert_same, _compare_world

pytestmark = pytest.mark.gpu


def test_world_without_bodies(pk, oracle):
    box = oracle.box_verts((0.5, 0.5, 0.5))
    s = {"shape_offsets": np.array([0, 8], np.uint32), "verts": box,
         "is_static": np.zeros(0, np.uint8), "shape": np.zeros(0, np.uint32),
         "pos": np.zeros((0, 3)), "vel": np.zeros((0, 3)), "angvel": np.zeros((0, 3)),
         "quat": np.zeros((0, 4)), "mass": np.zeros(0), "inertia": np.zeros((0, 9)),
         "restitution": np.zeros(0), "friction": np.zeros(0)}
    print("[edges] world without bodies", flush=True)
    g = pk.GpuWorld()
    g.load_scene(s)
    for _ in range(3):
        g.step(1 / 60)
        c = g.counters()
        assert c["pairs"] == 0 and c["gjk_hits"] == 0 and c["rows"] == 0
    assert len(g.pairs()) == 0 and len(g.bodies()["pos"]) == 0
    g.close()


def test_batched_partition_with_empty_and_single_body_worlds(pk, oracle):
    """Empty worlds between, before and after the others, and worlds of one body: the
    non-empty worlds step exactly as they do on their own."""
    s = pk.scene(pk.SCENE_BATCHED, 779, 5)
    fb = s["first_body"].astype(np.int64)
    n = int(fb[-1])
    # split off each world's first body as a world of its own, and put empty worlds around
    cuts = [0, 0]
    for w in range(len(fb) - 1):
        cuts += [int(fb[w]), int(fb[w]) + 1, int(fb[w]) + 1]
    cuts += [n, n]
    part = np.array(sorted(cuts), np.uint32)
    nw = len(part) - 1
    print(f"[edges] {nw} worlds, {int(np.sum(np.diff(part) == 0))} empty", flush=True)
    g = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"])
    g.load_scene(s)
    g.set_worlds(part)
    worlds = []
    for w in range(nw):
        b0, b1 = int(part[w]), int(part[w + 1])
        if b1 == b0:
            continue
        ws = bu.body_range(s, b0, b1)
        o = oracle.OracleWorld(iterations=ws["iterations"], solver_mode=1, flags=ws["flags"])
        o.load_scene(ws)
        worlds.append((o, b0, b1))
    for k in range(40):
        g.step(s["dt"])
        gb = g.bodies()
        gk = g.pairs()
        gc, gd = g.manifolds()
        for o, b0, b1 in worlds:
            o.step(s["dt"])
            ob = o.bodies()
            for f in ("pos", "vel", "angvel", "quat", "fat", "moved"):
                _assert_same((b0, k), f, gb[f][b0:b1], ob[f])
            local, m = bu.keys_in(gk, b0, b1)
            _assert_same((b0, k), "pairs", local, o.pairs())
            oc, od = o.manifolds()
            _assert_same((b0, k), "manifold counts", gc[m], oc)
            _assert_same((b0, k), "manifold data", gd[m], od)
    g.close()


def test_fast_bodies_bitexact(pk, oracle):
    """Boxes at about 15 m/s with no gravity: about half the leaves leave their fat boxes
    every step, so the pair list, manifolds and the refitted tree change wholesale
    between the full LBVH builds (EPA guard on: faster impacts drive the oracle's own
    run to non-finite values)."""
    s = pk.scene(pk.SCENE_BOXES, 31, 256)
    s = dict(s)
    rng = np.random.default_rng(3)
    s["vel"] = rng.normal(scale=15.0, size=s["vel"].shape) * (s["is_static"] == 0)[:, None]
    s["angvel"] = rng.normal(scale=5.0, size=s["angvel"].shape) * (s["is_static"] == 0)[:, None]
    flags = s["flags"] | pk.FLAG_EPA_GUARD
    o = oracle.OracleWorld(gravity=(0, 0, 0), iterations=s["iterations"], solver_mode=1, flags=flags)
    o.load_scene(s)
    g = pk.GpuWorld(gravity=(0, 0, 0), iterations=s["iterations"], flags=flags)
    g.load_scene(s)
    print(f"[edges] fast bodies: {len(s['shape'])}", flush=True)
    moved = hits = 0
    for k in range(60):
        g.step(s["dt"])
        o.step(s["dt"])
        _compare_world(k, g, o)
        c = o.counters()
        moved += c["moved"]
        hits += c["gjk_hits"]
    assert moved > 60 * 100 and hits > 0
    assert np.isfinite(g.bodies()["pos"]).all()
    g.close()


def test_drifting_hulls_record_cull_bitexact(pk, oracle):
    """Hull clusters (several pairs per body, so the cull reads per-body records and
    settles the culled pairs' manifolds) drifting apart without gravity: contacts form,
    then break as pairs get culled, and some steps move no leaf (the pair list is kept
    whole, with no source map)."""
    s = dict(pk.scene(pk.SCENE_HULLS, 5, 1024))
    rng = np.random.default_rng(11)
    s["vel"] = rng.normal(scale=1.0, size=s["vel"].shape)
    s["angvel"] = rng.normal(scale=1.0, size=s["angvel"].shape)
    flags = pk.FLAG_EPA_GUARD
    o = oracle.OracleWorld(gravity=(0, 0, 0), iterations=s["iterations"], solver_mode=1, flags=flags)
    o.load_scene(s)
    g = pk.GpuWorld(gravity=(0, 0, 0), iterations=s["iterations"], flags=flags)
    g.load_scene(s)
    n = len(s["shape"])
    print(f"[edges] drifting hulls: {n}", flush=True)
    still = record = hits = ends = 0
    for k in range(40):
        g.step(s["dt"])
        o.step(s["dt"])
        _compare_world(k, g, o)
        c = o.counters()
        still += c["moved"] == 0 and c["pairs"] > 0
        record += c["pairs"] >= 2 * n
        hits += c["gjk_hits"]
        ends += int(np.sum(g.events()[1] == 2))
    assert still >= 1 and record >= 30 and hits > 0 and ends > 0
    g.close()
