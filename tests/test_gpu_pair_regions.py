"""The pair query's multi-region output (taken above a leaf count, normally about a
million leaves): forced on small scenes through PKG_PAIR_REGION_MIN_LEAVES, which the
library reads when a context is created. Pair sets, manifolds and bodies stay bit-exact
against the oracle."""
import pytest

from test_gpu_parity import _run

pytestmark = pytest.mark.gpu


def test_pair_regions_hulls(pk, oracle, monkeypatch):
    monkeypatch.setenv("PKG_PAIR_REGION_MIN_LEAVES", "0")
    print("[regions] hulls 8192, 2 steps", flush=True)
    t = _run(pk, oracle, 1, 8192, 2)
    assert t["hits"] > 0


def test_pair_regions_boxes_world(pk, oracle, monkeypatch):
    monkeypatch.setenv("PKG_PAIR_REGION_MIN_LEAVES", "0")
    print("[regions] boxes 64, 30 steps", flush=True)
    _run(pk, oracle, 0, 64, 30)
