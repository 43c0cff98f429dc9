"""The BVH builder's rules restated in numpy, recursively (test infrastructure: the native builder,
``lumen_amd/csrc/bvh_build.cpp``, must build the same trees)."""

import numpy as np

LEAF_SIZE = 1
BINS = 16


def _area(lo, hi):
    e = np.maximum(hi - lo, 0.0)
    return 2.0 * (e[..., 0] * e[..., 1] + e[..., 1] * e[..., 2] + e[..., 2] * e[..., 0])


def _split(lo, hi, idx, bins):
    """Return (left, right) index arrays of the best binned SAH split of boxes ``idx``, or None (make a leaf)."""
    c = 0.5 * (lo[idx] + hi[idx])
    clo, chi = c.min(0), c.max(0)
    best = (np.inf, None)
    for ax in range(3):
        ext = chi[ax] - clo[ax]
        if ext <= 0.0:
            continue
        b = np.minimum(((c[:, ax] - clo[ax]) / ext * bins).astype(int), bins - 1)
        for k in range(1, bins):
            left = b < k
            nl = int(left.sum())
            if nl == 0 or nl == len(idx):
                continue
            li, ri = idx[left], idx[~left]
            cost = nl * _area(lo[li].min(0), hi[li].max(0)) + (len(idx) - nl) * _area(lo[ri].min(0), hi[ri].max(0))
            if cost < best[0]:
                best = (cost, (li, ri))
    if best[1] is None:  # every centroid coincides: halve the list
        h = len(idx) // 2
        return idx[:h], idx[h:]
    return best[1]


def build_py(lo: np.ndarray, hi: np.ndarray, leaf_size: int = LEAF_SIZE, bins: int = BINS) -> dict:
    """Return one tree over boxes (lo, hi) (n, 3): dict(lo, hi, skip, start, count, items, depth).

    ``skip`` and ``start`` are relative to this tree (node 0 and item 0); ``items`` is a permutation of range(n).
    """
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    n = len(lo)
    if n == 0:
        raise ValueError("bvh.build: no boxes")
    nodes_lo, nodes_hi, skip, start, count, items = [], [], [], [], [], []
    depth = 0

    def rec(idx, d):
        nonlocal depth
        depth = max(depth, d)
        me = len(nodes_lo)
        nodes_lo.append(lo[idx].min(0))
        nodes_hi.append(hi[idx].max(0))
        skip.append(0)
        start.append(0)
        count.append(0)
        if len(idx) <= leaf_size:
            start[me] = len(items)
            count[me] = len(idx)
            items.extend(int(i) for i in idx)
        else:
            li, ri = _split(lo, hi, idx, bins)
            rec(li, d + 1)
            rec(ri, d + 1)
        skip[me] = len(nodes_lo)

    rec(np.arange(n), 1)
    return dict(lo=np.asarray(nodes_lo), hi=np.asarray(nodes_hi), skip=np.asarray(skip, np.int32),
                start=np.asarray(start, np.int32), count=np.asarray(count, np.int32),
                items=np.asarray(items, np.int32), depth=depth)


