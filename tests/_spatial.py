"""The draw order of the differentiable path's spatial processing mode, restated on the CPU (test
helper): splats of a scene rendered by four or more views of one call are processed in Morton order
of their means (10 bits per axis, quantized in the scene's bounding box, ties by index), so visible
splats of equal depth key blend in that order instead of ascending id."""
import numpy as np

import oracle


def _spread3(x):
    x = x.astype(np.uint32)
    x = (x | (x << 16)) & np.uint32(0x030000FF)
    x = (x | (x << 8)) & np.uint32(0x0300F00F)
    x = (x | (x << 4)) & np.uint32(0x030C30C3)
    x = (x | (x << 2)) & np.uint32(0x09249249)
    return x


def morton_rank(means, axis_bits=10):
    """rank[id] = position of splat id in the Morton processing order."""
    m = np.ascontiguousarray(means, np.float32)
    top = np.float32((1 << axis_bits) - 1)
    code = np.zeros(m.shape[0], np.uint32)
    for a in range(3):
        x = m[:, a]
        lo, hi = x.min(), x.max()
        if hi > lo:
            t = (x - lo) / (hi - lo)
        else:
            t = np.zeros_like(x)
        qf = np.minimum(np.maximum(t * top, np.float32(0)), top)
        code |= _spread3(qf.astype(np.uint32)) << np.uint32(a)
    perm = np.argsort(code, kind="stable")
    rank = np.empty_like(perm)
    rank[perm] = np.arange(perm.size)
    return rank


def draw_order(cam, pos, rank=None):
    """Splat ids in draw order for one view: visible splats by ascending depth key, equal keys by
    ascending id (rank None) or by Morton rank."""
    _, key, index = oracle.rank(cam, pos)
    if rank is None:
        _, order = oracle.sort_kv(key, index)
        return order.astype(np.int64)
    o = np.lexsort((rank[index], key))
    return index[o].astype(np.int64)
