"""Write tests/golden/nav_entry_example_tank/: the example tank's navigator arrays as a cache entry in the reference's
layout (geometry/cache.py:187-230 ``_encode``/``_save``; geometry/navigator_build.py:51-90 the NavArrays fields and
their order, :371-397 the host arrays with their mesh placeholders, :401-420 the surface-code layout), encoded here by
a restatement of those rules of its own (not lumen_amd's writer), so the fixture checks the writer and the reader.

    python tests/golden/make_nav_entry.py
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def encode(x, arrays):
    """The reference's _encode rules: arrays by key a<i>, dataclass records by field, dict/tuple/list, scalars."""
    if isinstance(x, (np.ndarray, np.generic)):
        key = "a%d" % len(arrays)
        arrays[key] = np.asarray(x)
        return {"a": key, "scalar": isinstance(x, np.generic)}
    if isinstance(x, tuple) and len(x) == 3 and x[0] == "__record__":
        return {"dc": x[1], "f": {k: encode(v, arrays) for k, v in x[2]}}
    if isinstance(x, dict):
        return {"m": [[encode(k, arrays), encode(v, arrays)] for k, v in x.items()]}
    if isinstance(x, tuple):
        return {"t": [encode(v, arrays) for v in x]}
    if isinstance(x, list):
        return {"l": [encode(v, arrays) for v in x]}
    assert x is None or isinstance(x, (bool, int, float, str)), type(x)
    return x


def main():
    from lumen_amd.detector import example_tank

    nav = example_tank().nav
    n = int(nav.n)
    host = dict(nav.host)
    for k in ("V0", "V1", "V2", "N"):
        host[k] = np.zeros((0, 3), np.float32)
    for k in ("cstart", "cn", "cmesh", "tchunk"):
        host[k] = np.zeros(1, np.int32)
    T = nav.csg["T"]
    fm = 1 + int(max(T["np_"]))
    su = n + 1
    sm = su + int(host["ub"][-1]) * fm * 4
    parent = [int(p) for p in host["parent"][:n]]
    kids = tuple(tuple(i for i in range(n) if parent[i] == v) for v in range(n + 1))
    csg = ("__record__", "lumen.geometry.csg_build:CsgTables",
           [("T", dict(T)), ("K", dict(nav.csg["K"])), ("G", dict(nav.csg["G"])), ("kind_box", ()),
            ("empty_leaf", int(nav.csg["empty_leaf"])), ("has_torus", bool(np.any(T["typ"] == 2)))])
    rec = ("__record__", "lumen.geometry.navigator_build:NavArrays",
           [("host", host), ("csg", csg), ("n", n), ("n_mesh", 0), ("chunk", 8), ("ch0", n + 1), ("srf", (su, sm, fm)),
            ("names", tuple(nav.names)), ("defs", ()), ("stats", dict(nav.stats)), ("owned", ()), ("clipped", ()),
            ("kids", kids), ("lo_w", np.asarray(nav.lo_w, np.float64)), ("hi_w", np.asarray(nav.hi_w, np.float64)),
            ("probe", {}), ("origins", tuple(nav.names) + ("OUT",)),
            ("pieces", tuple(int(x) for x in host["piece"][:n])), ("thin", ()), ("precision", "float32")])
    arrays = {}
    tree = encode(rec, arrays)
    out = os.path.join(HERE, "nav_entry_example_tank")
    os.makedirs(out, exist_ok=True)
    np.savez(os.path.join(out, "arrays.npz"), **arrays)
    with open(os.path.join(out, "meta.json"), "w") as f:
        json.dump(tree, f)
    print(out, len(arrays), "arrays")


if __name__ == "__main__":
    main()
