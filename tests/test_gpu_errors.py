"""Error paths of the C ABI that need a device context (pkg_create refuses to run
without a GPU): uploads that would leave bodies pointing at missing shapes."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_shape_upload_cannot_drop_a_shape_in_use(pk):
    s = pk.scene(pk.SCENE_BOXES, None, 64)
    g = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"])
    g.load_scene(s)
    used = int(s["shape"].max())
    offs = np.ascontiguousarray(s["shape_offsets"][:used + 1], dtype=np.uint32)  # drops shape `used`
    verts = np.ascontiguousarray(s["verts"][:offs[-1]], dtype=np.float64)
    st = g.L.pkg_upload_shapes(g.h, len(offs) - 1, offs.ctypes.data_as(C.POINTER(C.c_uint32)),
                               verts.ctypes.data_as(C.POINTER(C.c_double)))
    assert st == 1, st
    assert b"upload the bodies again" in g.L.pkg_last_error(g.h)
    # the same shapes again are accepted, and the world still steps bit-identically to a fresh one
    offs = np.ascontiguousarray(s["shape_offsets"], dtype=np.uint32)
    verts = np.ascontiguousarray(s["verts"], dtype=np.float64)
    assert g.L.pkg_upload_shapes(g.h, len(offs) - 1, offs.ctypes.data_as(C.POINTER(C.c_uint32)),
                                 verts.ctypes.data_as(C.POINTER(C.c_double))) == 0
    h = pk.GpuWorld(iterations=s["iterations"], flags=s["flags"])
    h.load_scene(s)
    for _ in range(10):
        g.step(s["dt"])
        h.step(s["dt"])
    np.testing.assert_array_equal(g.bodies()["pos"], h.bodies()["pos"])
    g.close()
    h.close()
