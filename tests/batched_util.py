"""Per-world views of a batched scene (bodies [first_body[w], first_body[w+1]) form
world w, with its shapes contiguous) for checking each world against its own oracle."""
import numpy as np


def world_slice(s, w):
    fb = s["first_body"]
    b0, b1 = int(fb[w]), int(fb[w + 1])
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
    return out, b0, b1


def keys_in(keys, b0, b1):
    """Keys of world [b0, b1) renumbered to the world's own ids, and the row mask."""
    a = (keys >> np.uint64(32)).astype(np.int64)
    b = (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    m = (a >= b0) & (a < b1)
    local = ((a[m] - b0).astype(np.uint64) << np.uint64(32)) | (b[m] - b0).astype(np.uint64)
    return local, m
