"""Several steps in one call: pkg_step(ctx, dt, k) and pkg_step_timed run each step's
integrate_pos fused with the next step's integrate_vel (one kernel) and return with
every step integrated. The result must be bit-identical to k calls of one step, and to
the oracle stepped k times; the running counter totals add up the steps' own counts."""
import numpy as np
import pytest

from test_gpu_parity import _assert_same, _compare_world

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "angvel", "quat", "fat", "moved")


def _same_worlds(tag, a, b):
    ba, bb = a.bodies(), b.bodies()
    for f in FIELDS:
        _assert_same(tag, f, ba[f], bb[f])
    _assert_same(tag, "pairs", a.pairs(), b.pairs())
    ca, da = a.manifolds()
    cb, db = b.manifolds()
    _assert_same(tag, "manifold counts", ca, cb)
    _assert_same(tag, "manifold data", da, db)


def test_multistep_calls_match_single_steps_and_oracle(pk, oracle):
    """Boxes falling onto the ground and stacking (contacts from about step 20): calls of
    1, 7, 12 and 10 steps, against 30 single steps and the oracle."""
    s = pk.scene(pk.SCENE_BOXES, 41, 256)
    one = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"])
    one.load_scene(s)
    many = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"])
    many.load_scene(s)
    o = oracle.OracleWorld(iterations=s["iterations"], solver_mode=1, flags=s["flags"])
    o.load_scene(s)
    done = 0
    pairs = hits = 0
    for k in (1, 7, 12, 10, 30):
        print(f"[multistep] call of {k} steps after {done}", flush=True)
        many.step(s["dt"], k)
        for _ in range(k):
            one.step(s["dt"])
            c = one.counters()
            pairs += c["pairs"]
            hits += c["gjk_hits"]
            o.step(s["dt"])
        done += k
        _same_worlds(done, many, one)
        _compare_world(done, many, o)
        c = many.counters()
        assert c["steps_total"] == done
        assert c["pairs_total"] == pairs and c["gjk_hits_total"] == hits
    assert hits > 0
    one.close()
    many.close()


def test_timed_multistep_matches_single_steps_batched(pk, oracle):
    """pkg_step_timed over 15 steps of a few batched worlds in the mixed form against 15
    single steps: bit-identical, and every stage timed."""
    s = pk.scene(pk.SCENE_BATCHED, 779, 6)
    flags = s["flags"] | pk.FLAG_MIXED
    worlds = []
    for _ in range(2):
        g = pk.GpuWorld(iterations=s["iterations"], flags=flags)
        g.load_scene(s)
        g.set_worlds(s["first_body"])
        worlds.append(g)
    a, b = worlds
    print(f"[multistep] batched timed, {len(s['shape'])} bodies", flush=True)
    for _ in range(3):
        a.step(s["dt"])
        b.step(s["dt"])
    ms = a.step_timed(s["dt"], 15)
    for _ in range(15):
        b.step(s["dt"])
    _same_worlds("timed", a, b)
    assert ms["integrate_vel"] > 0 and ms["integrate_pos"] > 0 and ms["gjk"] > 0
    assert a.counters()["steps_total"] == b.counters()["steps_total"] == 18
    a.close()
    b.close()
