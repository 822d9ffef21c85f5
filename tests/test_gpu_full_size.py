"""GPU parity at the benchmark configurations' full sizes, bit for bit against the oracle:
  * the narrowphase stress scene, 262,144 convex hulls, two steps (every leaf moved,
    no solver): about a million pairs per step;
  * the 4,096 batched uneven worlds (2.97 million bodies), three steps, each world
    against its own oracle world (worlds are stepped on a pool of host threads; the
    oracle releases the interpreter during a step). Worlds are contiguous body ranges,
    so the concatenation of the oracle worlds' sorted pair lists, shifted to global ids,
    is the GPU's sorted list: the comparison runs on whole arrays.
Both run the FP64 drop-in and the mixed-precision filters (FP32 leaf tests and GJK
separation, which must not change a result; the solve stays FP64 here)."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import batched_util as bu
from test_gpu_parity import _assert_same

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("filters", [False, True], ids=["fp64", "fp32_filters"])
def test_hulls_full_size(pk, oracle, filters):
    s = pk.scene(pk.SCENE_HULLS, None, 0)
    flags = s["flags"] | ((pk.FLAG_FP32_GJK | pk.FLAG_FP32_BROAD) if filters else 0)
    print(f"[full] hulls n={len(s['shape'])} filters={filters}", flush=True)
    o = oracle.OracleWorld(iterations=s["iterations"], solver_mode=1, flags=s["flags"])
    o.load_scene(s)
    g = pk.GpuWorld(iterations=s["iterations"], flags=flags)
    g.load_scene(s)
    for k in range(2):
        o.step(s["dt"])
        g.step(s["dt"])
        _assert_same(k, "pairs", g.pairs(), o.pairs())
        gc, gd = g.manifolds()
        oc, od = o.manifolds()
        _assert_same(k, "manifold counts", gc, oc)
        _assert_same(k, "manifold data", gd, od)
        assert g.counters()["gjk_hits"] == o.counters()["gjk_hits"]
        print(f"[full] hulls step {k}: pairs={len(oc)} hits={o.counters()['gjk_hits']}", flush=True)
    g.close()


@pytest.mark.parametrize("filters", [False, True], ids=["fp64", "fp32_filters"])
def test_batched_worlds_full_size(pk, oracle, filters):
    s = pk.scene(pk.SCENE_BATCHED, None, 0)
    fb = s["first_body"]
    nw = len(fb) - 1
    flags = s["flags"] | ((pk.FLAG_FP32_GJK | pk.FLAG_FP32_BROAD) if filters else 0)
    print(f"[full] batched worlds={nw} bodies={len(s['shape'])} filters={filters}", flush=True)
    g = pk.GpuWorld(iterations=s["iterations"], flags=flags)
    g.load_scene(s)
    g.set_worlds(fb)

    def make(w):
        ws, b0, b1 = bu.world_slice(s, w)
        o = oracle.OracleWorld(iterations=ws["iterations"], solver_mode=1, flags=ws["flags"])
        o.load_scene(ws)
        return o

    threads = max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(threads) as ex:
        worlds = list(ex.map(make, range(nw)))
        print(f"[full] {nw} oracle worlds loaded", flush=True)
        for k in range(3):
            g.step(s["dt"])
            list(ex.map(lambda o: o.step(s["dt"]), worlds))
            obs = [o.bodies() for o in worlds]
            gb = g.bodies()
            for f in ("pos", "vel", "angvel", "quat", "fat", "moved"):
                _assert_same(k, f, gb[f], np.concatenate([ob[f] for ob in obs]))
            keys = []
            for w, o in enumerate(worlds):
                p = o.pairs()
                shift = np.uint64(int(fb[w]))
                keys.append(((p >> np.uint64(32)) + shift) << np.uint64(32) | ((p & np.uint64(0xFFFFFFFF)) + shift))
            _assert_same(k, "pairs", g.pairs(), np.concatenate(keys))
            gc, gd = g.manifolds()
            mf = [o.manifolds() for o in worlds]
            _assert_same(k, "manifold counts", gc, np.concatenate([m[0] for m in mf]))
            _assert_same(k, "manifold data", gd, np.concatenate([m[1] for m in mf]))
            rows = sum(o.counters()["rows"] for o in worlds)
            assert g.counters()["rows"] == rows
            print(f"[full] batched step {k}: pairs={len(gc)} rows={rows}", flush=True)
    g.close()
