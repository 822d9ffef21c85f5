"""pkg_gjk_epa_batch_device, the entry point that allocates no device memory, against
pkg_gjk_epa_batch (the allocating entry point it is a variant of) on the same pairs:
hit codes and contact data bit for bit, on two calls in a row, the second smaller, so
the static scratch left by the first call cannot leak into the second. Inputs and
outputs are device arrays the test allocates itself (the caller's part)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree through ctypes, from the HIP runtime the library
    already loaded."""

    def __init__(self):
        path = "/opt/rocm/lib/libamdhip64.so"
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line:
                    path = line.split()[-1]
                    break
        self.L = C.CDLL(path)
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.live.append(p)
        if a.nbytes:
            assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def get(self, p, like):
        out = np.zeros_like(like)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


def pairs(oracle, n, seed):
    """n shape pairs of every kind (AABBs, OBBs, box, hull and UV-sphere meshes), about
    half of them overlapping, in the layout of pkg_gjk_epa_batch."""
    rng = np.random.default_rng(seed)
    meshes = [oracle.box_verts((0.4, 0.3, 0.5)), oracle.sphere_verts(0.5, 16, 32),
              oracle.sphere_verts(0.35, 32, 64), rng.normal(size=(20, 3)) * 0.3, oracle.pyramid_verts(0.5, 0.7)]
    offs = np.cumsum([0] + [len(m) for m in meshes])
    verts = np.concatenate(meshes)
    kind = np.zeros((2, n), np.int32)
    prm = np.zeros((2, n, 10))
    vb = np.zeros((2, n), np.uint32)
    ve = np.zeros((2, n), np.uint32)
    for i in range(n):
        for side in range(2):
            k = int(rng.integers(0, 3)) if rng.random() < 0.3 else 2
            q = rng.normal(size=4)
            q /= np.linalg.norm(q)
            c = rng.normal(size=3) * 0.35 + (0.0 if side == 0 else 0.4)
            kind[side, i] = k
            if k == 0:
                h = rng.uniform(0.1, 0.6, 3)
                prm[side, i, :6] = np.concatenate([c - h, c + h])
            elif k == 1:
                prm[side, i] = np.concatenate([c, q, rng.uniform(0.1, 0.6, 3)])
            else:
                m = int(rng.integers(0, len(meshes)))
                prm[side, i, :7] = np.concatenate([c, q])
                vb[side, i], ve[side, i] = offs[m], offs[m + 1]
    return kind, prm, vb, ve, verts


def call_host(pk, g, kind, prm, vb, ve, verts):
    n = kind.shape[1]
    out = np.zeros((n, 10))
    hit = np.zeros(n, np.int32)
    U32, I32, D = C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.POINTER(C.c_double)
    arrs = [np.ascontiguousarray(x) for x in (kind[0], prm[0], vb[0], ve[0], kind[1], prm[1], vb[1], ve[1], verts)]
    st = g.L.pkg_gjk_epa_batch(g.h, n, arrs[0].ctypes.data_as(I32), arrs[1].ctypes.data_as(D),
                               arrs[2].ctypes.data_as(U32), arrs[3].ctypes.data_as(U32),
                               arrs[4].ctypes.data_as(I32), arrs[5].ctypes.data_as(D),
                               arrs[6].ctypes.data_as(U32), arrs[7].ctypes.data_as(U32),
                               arrs[8].ctypes.data_as(D), len(verts), out.ctypes.data_as(D),
                               hit.ctypes.data_as(I32))
    assert st == 0, g.L.pkg_last_error(g.h)
    return out, hit


def call_device(g, hip, kind, prm, vb, ve, verts, out_fill=0.0):
    n = kind.shape[1]
    d = [hip.put(x) for x in (kind[0], prm[0], vb[0], ve[0], kind[1], prm[1], vb[1], ve[1], verts)]
    out_h = np.full((n, 10), out_fill)
    hit_h = np.full(n, 99, np.int32)
    d_out, d_hit = hip.put(out_h), hip.put(hit_h)
    fn = g.L.pkg_gjk_epa_batch_device
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_uint32] + [C.c_void_p] * 8 + [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
    st = fn(g.h, n, *d[:8], d[8], len(verts), d_out, d_hit)
    return st, hip.get(d_out, out_h), hip.get(d_hit, hit_h)


def test_no_alloc_batch_matches_allocating_batch(pk, oracle):
    g = pk.GpuWorld()
    hip = Hip()
    try:
        for call, (n, seed) in enumerate(((6000, 5), (700, 6))):
            kind, prm, vb, ve, verts = pairs(oracle, n, seed)
            print(f"[no_alloc] call {call}: n={n} seed={seed}", flush=True)
            ref_out, ref_hit = call_host(pk, g, kind, prm, vb, ve, verts)
            st, out, hit = call_device(g, hip, kind, prm, vb, ve, verts, out_fill=-7.0)
            assert st == 0, g.L.pkg_last_error(g.h)
            np.testing.assert_array_equal(hit, ref_hit)
            m = ref_hit == 1
            assert m.sum() > n // 5
            np.testing.assert_array_equal(out[m], ref_out[m])
            hip.free()
    finally:
        hip.free()
        g.close()


def test_no_alloc_batch_rejects_what_it_cannot_hold(pk, oracle):
    g = pk.GpuWorld()
    hip = Hip()
    try:
        fn = g.L.pkg_gjk_epa_batch_device
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p, C.c_uint32] + [C.c_void_p] * 8 + [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
        # above the scratch's capacity: refused before any device work
        assert fn(g.h, 65537, *([None] * 8), None, 0, None, None) == 1
        assert b"PKG_NOALLOC_MAX_PAIRS" in g.L.pkg_last_error(g.h)
        # a pair with a kind out of range, and one with a vertex range past n_verts
        kind, prm, vb, ve, verts = pairs(oracle, 40, 8)
        kind[0, 3] = 7
        kind[1, 9] = 2
        vb[1, 9], ve[1, 9] = len(verts) - 2, len(verts) + 5
        st, out, hit = call_device(g, hip, kind, prm, vb, ve, verts)
        assert st == 1
        assert hit[3] == -1 and hit[9] == -1
        assert set(np.unique(np.delete(hit, [3, 9]))) <= {0, 1, 2}
    finally:
        hip.free()
        g.close()
