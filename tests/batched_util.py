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
