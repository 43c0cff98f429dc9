"""Bounding-volume hierarchies over daughter boxes: binned SAH, stored depth first with skip links.

The stored form is the reference navigator-array format's (``geometry/bvh.py:1-10``, ``:179-249``): nodes in depth
first order, a node's left child right after it, ``skip`` the node after its subtree (the stackless walk), a leaf
(``count`` > 0) covering ``items[start:start + count]``.  The builder is native (``csrc/bvh_build.cpp``, the C ABI's
``lh_bvh_build``): binned surface-area-heuristic splits (16 bins per axis, the cheapest plane over the three axes, a
stable partition), ``leaf_size`` boxes per leaf.
"""

from __future__ import annotations

import numpy as np

# one box per leaf: for rays from inside a container to daughters on its walls it visits as many nodes as two
# per leaf and tests half the daughters (the costly part of a walk)
LEAF_SIZE = 1
BINS = 16


def build(lo: np.ndarray, hi: np.ndarray, leaf_size: int = LEAF_SIZE, bins: int = BINS) -> dict:
    """Return one tree over boxes (lo, hi) (n, 3): dict(lo, hi, skip, start, count, items, depth), built by the
    native builder (``lh_bvh_build``, ``csrc/bvh_build.cpp``).

    ``skip`` and ``start`` are relative to this tree (node 0 and item 0); ``items`` is a permutation of range(n).
    """  # noqa: D205
    from .. import _lib

    lo = np.ascontiguousarray(lo, np.float64).reshape(-1, 3)
    hi = np.ascontiguousarray(hi, np.float64).reshape(-1, 3)
    n = len(lo)
    if n == 0:
        raise ValueError("bvh.build: no boxes")
    m = 2 * n
    nlo, nhi = np.zeros((m, 3)), np.zeros((m, 3))
    skip, start, count = (np.zeros(m, np.int32) for _ in range(3))
    items = np.zeros(n, np.int32)
    nn, depth = _lib.C.c_int32(), _lib.C.c_int32()
    p = lambda a: a.ctypes.data_as(_lib.C.c_void_p)  # noqa: E731
    _lib.check(_lib.lib().lh_bvh_build(p(lo), p(hi), n, int(leaf_size), int(bins), p(nlo), p(nhi), p(skip), p(start),
                                        p(count), p(items), _lib.C.byref(nn), _lib.C.byref(depth)), "lh_bvh_build")
    k = nn.value
    return dict(lo=nlo[:k], hi=nhi[:k], skip=skip[:k], start=start[:k], count=count[:k], items=items,
                depth=depth.value)


def slab(lo, hi, o, d):
    """Return (entry, exit) of the ray o + t d through the box [lo, hi] (float64, host; d without zero components)."""
    inv = 1.0 / d
    ta, tb = (lo - o) * inv, (hi - o) * inv
    return np.max(np.minimum(ta, tb)), np.min(np.maximum(ta, tb))
