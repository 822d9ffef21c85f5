"""GPU parity of the mixed-precision form (PKG_FLAG_MIXED): FP32 broadphase leaf tests
rounded outward, FP32 certified GJK separation and the FP32 contact solve, with EPA,
manifolds and body state in FP64.

  * The two FP32 filters never change a result: with the FP64 solve they leave every
    body, pair, manifold, event and colour bit-exact against the oracle over whole
    runs (batched uneven worlds and the hull clusters).
  * With the FP32 solve as well, each step is checked from a shared snapshot: the
    GPU's state after step k is loaded into the oracle (bodies, leaf bounds, pairs and
    manifolds), both take step k+1, and then the pair set, GJK hit count and the
    manifolds' geometry (normals, contact points, depths) are bit-exact, the
    accumulated impulses close, and every position within 1e-5 relative to
    max(|x|, 1 m) of the oracle's FP64 result."""
import numpy as np
import pytest

import batched_util as bu
from test_gpu_parity import _assert_same, _run

pytestmark = pytest.mark.gpu


def test_fp32_filters_bitexact_batched_worlds(pk, oracle):
    s = pk.scene(pk.SCENE_BATCHED, 778, 6)
    fb = s["first_body"]
    flags = s["flags"] | pk.FLAG_FP32_GJK | pk.FLAG_FP32_BROAD
    g = pk.GpuWorld(iterations=s["iterations"], flags=flags)
    g.load_scene(s)
    g.set_worlds(fb)
    worlds = []
    for w in range(len(fb) - 1):
        ws, b0, b1 = bu.world_slice(s, w)
        o = oracle.OracleWorld(iterations=ws["iterations"], solver_mode=1, flags=ws["flags"])
        o.load_scene(ws)
        worlds.append((o, b0, b1))
    print(f"[mixed] fp32 filters, batched worlds sizes={np.diff(fb).tolist()}", flush=True)
    listed = fp64 = rows = 0
    for k in range(60):
        g.step(s["dt"])
        c = g.counters()
        listed += c["gjk_listed"]
        fp64 += c["gjk_fp64"]
        gb, gk = g.bodies(), g.pairs()
        gc, gd = g.manifolds()
        for w, (o, b0, b1) in enumerate(worlds):
            o.step(s["dt"])
            ob = o.bodies()
            for f in ("pos", "vel", "angvel", "quat", "fat", "moved"):
                _assert_same((w, k), f, gb[f][b0:b1], ob[f])
            local, m = bu.keys_in(gk, b0, b1)
            _assert_same((w, k), "pairs", local, o.pairs())
            oc, od = o.manifolds()
            _assert_same((w, k), "manifold counts", gc[m], oc)
            _assert_same((w, k), "manifold data", gd[m], od)
            rows += o.counters()["rows"]
    g.close()
    print(f"[mixed] GJK stage: {listed} pairs listed, {fp64} ran the FP64 GJK", flush=True)
    assert rows > 0
    assert fp64 < listed  # the FP32 pass certified some pairs apart


def test_fp32_filters_bitexact_hulls(pk, oracle):
    t = _run(pk, oracle, 1, 8192, 2,
             flags=pk.scene(1, None, 8)["flags"] | pk.FLAG_FP32_GJK | pk.FLAG_FP32_BROAD)
    assert t["hits"] > 0


def test_fp32_filters_bitexact_granular(pk, oracle):
    """The granular pile (about 100 pairs per body: the cull reads per-body records and,
    in the mixed form, tests FP32 spheres first) with the FP32 filters: bit-exact."""
    s = pk.scene(4, None, 343)
    t = _run(pk, oracle, 4, 343, 60, flags=s["flags"] | pk.FLAG_FP32_GJK | pk.FLAG_FP32_BROAD)
    assert t["pairs"] > 2 * 343 * 60


def test_fp32_filters_bitexact_boxes(pk, oracle):
    t = _run(pk, oracle, 0, 256, 60, flags=pk.FLAG_FP32_GJK | pk.FLAG_FP32_BROAD)
    assert t["rows"] > 0


def test_mixed_form_each_step_against_fp64_oracle(pk, oracle):
    s = pk.scene(pk.SCENE_BATCHED, 779, 5)
    fb = s["first_body"]
    g = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"] | pk.FLAG_MIXED)
    g.load_scene(s)
    g.set_worlds(fb)
    worlds = []
    for w in range(len(fb) - 1):
        ws, b0, b1 = bu.world_slice(s, w)
        o = oracle.OracleWorld(iterations=ws["iterations"], solver_mode=1, flags=ws["flags"])
        o.load_scene(ws)
        worlds.append((o, b0, b1))
    print(f"[mixed] full mixed form, batched worlds sizes={np.diff(fb).tolist()}", flush=True)
    worst = 0.0
    rows = 0
    for k in range(40):
        # snapshot of the GPU world after step k, loaded into each world's oracle
        gb, gk = g.bodies(), g.pairs()
        gc, gd = g.manifolds()
        for o, b0, b1 in worlds:
            local, m = bu.keys_in(gk, b0, b1)
            o.set_state({f: gb[f][b0:b1] for f in ("pos", "vel", "angvel", "quat", "fat")}, local, gc[m], gd[m])
        g.step(s["dt"])
        gb, gk = g.bodies(), g.pairs()
        gc, gd = g.manifolds()
        hits = 0
        for w, (o, b0, b1) in enumerate(worlds):
            o.step(s["dt"])
            ob = o.bodies()
            for f in ("fat", "moved"):
                _assert_same((w, k), f, gb[f][b0:b1], ob[f])
            local, m = bu.keys_in(gk, b0, b1)
            _assert_same((w, k), "pairs", local, o.pairs())
            oc, od = o.manifolds()
            _assert_same((w, k), "manifold counts", gc[m], oc)
            _assert_same((w, k), "manifold geometry", gd[m][:, :, :10], od[:, :, :10])
            scale = max(1.0, float(np.abs(od[:, :, 10:]).max(initial=0.0)))
            np.testing.assert_allclose(gd[m][:, :, 10:], od[:, :, 10:], rtol=1e-3, atol=1e-4 * scale)
            hits += o.counters()["gjk_hits"]
            rows += o.counters()["rows"]
            x64 = ob["pos"]
            err = float(np.max(np.abs(gb["pos"][b0:b1] - x64) / np.maximum(np.abs(x64), 1.0)))
            worst = max(worst, err)
        assert hits == g.counters()["gjk_hits"], (k, hits, g.counters()["gjk_hits"])
    g.close()
    print(f"[mixed] worst per-step position error vs FP64: {worst:.3e}", flush=True)
    assert rows > 0
    assert worst <= 1e-5
