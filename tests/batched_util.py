"""Per-world views of a batched scene for checking each world against its own oracle."""
import numpy as np

from physkit_amd import scene_world as world_slice  # noqa: F401  (re-exported for the tests)


def keys_in(keys, b0, b1):
    """Keys of world [b0, b1) renumbered to the world's own ids, and the row mask."""
    a = (keys >> np.uint64(32)).astype(np.int64)
    b = (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    m = (a >= b0) & (a < b1)
    local = ((a[m] - b0).astype(np.uint64) << np.uint64(32)) | (b[m] - b0).astype(np.uint64)
    return local, m


def body_range(s, b0, b1):
    """Bodies [b0, b1) of a scene as a scene of their own (renumbered from 0, with the
    contiguous shape range they use)."""
    shp = s["shape"][b0:b1].astype(np.int64)
    s0, s1 = int(shp.min()), int(shp.max()) + 1
    offs = s["shape_offsets"].astype(np.int64)
    out = {k: s[k][b0:b1].copy() for k in ("is_static", "pos", "vel", "angvel", "quat", "mass", "inertia",
                                            "restitution", "friction")}
    out["shape"] = (shp - s0).astype(np.uint32)
    out["shape_offsets"] = (offs[s0:s1 + 1] - offs[s0]).astype(np.uint32)
    out["verts"] = s["verts"][offs[s0]:offs[s1]].copy()
    for k in ("dt", "iterations", "steps", "flags"):
        out[k] = s[k]
    out["n_static"] = int(np.sum(out["is_static"] != 0))
    out["first_body"] = np.array([0, b1 - b0], np.uint32)
    return out
