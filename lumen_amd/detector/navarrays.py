"""The navigator-array format: volumes, daughter BVHs in frame groups and CSG tables, as named host arrays.

This is the reference's persistent navigator format (``geometry/navigator_build.py:51-90``, ``:371-397``,
``geometry/csg_build.py:31-46``, ``:165-210``, on disk ``geometry/cache.py:187-264``): a dict ``host`` of per-node,
per-group, per-navigator-id and per-volume arrays, and ``csg`` = (T, K, G) leaf, kind and program tables with the
reference's column names.  ``save``/``load`` write and read a cache entry in the reference's layout (``arrays.npz``
plus ``meta.json``, the reference's ``_encode`` tree).

The builder here covers the solids this engine navigates (quadric-and-plane leaves; no meshes, no tori), volumes that
lie inside their mother (no overlaps, so no ownership links), and daughter groups in the world frame or, for daughters
that share a rotation their world boxes fit badly, in that rotation's frame (``frame_groups``, ``gR``).
Navigator ids: volumes 0..n-1, the outside n.  ``R``, ``t`` map a volume's frame to the world: x_w = R x + t.

Two arrays are this build's own, accelerations the reference's format does not hold (absent, nothing changes but
the speed): ``host["skin"]`` (n+1,), per volume the CSG leaf of its daughter-free core, or -1.  A volume whose solid
is one convex leaf (a box, a full tube, an orb) and whose daughters all sit near its boundary (sensors on a tank's
walls) gets the leaf shrunk by the depth D of its deepest daughter plus a margin: no daughter's box meets the core,
so a ray skips every BVH box over the part of its path inside the core.  ``skin_core`` says how D is found.
``host["touch"]`` (n+1,), 1 where a sibling's box lies within the navigator's relocation step of the volume's box
(``touch_flags``): only such a volume's exits look for a sibling holding the point past the face.
"""

from __future__ import annotations

import getpass
import hashlib
import json
import os
import shutil
import tempfile
from dataclasses import dataclass, field

import numpy as np

from . import bvh
from .solids import Leaf, Solid, program_depth

HUGE = 1e30
BOX_PAD_REL = 1e-6
BOX_PAD_ABS = 1e-9
STACK = 32
MAX_DEPTH = 31


@dataclass
class Volume:
    """A placed volume: its solid (kind index), world pose, mother (-1: the outside), class and name."""

    name: str
    kind: int
    R: np.ndarray
    t: np.ndarray
    parent: int
    cls: int


@dataclass
class NavArrays:
    """The navigator's arrays (module docstring).  ``host``: named arrays; ``csg``: {"T", "K", "G"} tables."""

    host: dict
    csg: dict
    n: int
    names: tuple
    lo_w: np.ndarray
    hi_w: np.ndarray
    n_mesh: int = 0
    precision: str = "float32"
    stats: dict = field(default_factory=dict)

    # ------------------------------------------------------------------ on disk: arrays.npz + meta.json
    def save(self, path: str) -> None:
        """Write the arrays into directory ``path`` in the reference's cache entry layout (``geometry/cache.py:187-230``
        ``_encode``/``_save``): ``meta.json`` the tree of a ``NavArrays`` record (``dc`` the class, ``f`` its fields
        in the reference's order), its arrays in ``arrays.npz`` as ``a0``, ``a1``, ... in the order the tree meets
        them.  Fields this build does not hold are written empty: no meshes (``n_mesh`` 0, the mesh arrays one
        placeholder row), no ownership links, no definitions or placement-check probes."""  # noqa: D205, D209
        os.makedirs(path, exist_ok=True)
        store: dict = {}
        tree = _tree_of(self.reference_record(), store)
        np.savez(os.path.join(path, "arrays.npz"), **store)
        with open(os.path.join(path, "meta.json"), "w") as f:
            json.dump(tree, f)

    def reference_record(self) -> "_Record":
        """Return the arrays as the reference's ``NavArrays`` record (``geometry/navigator_build.py:51-90``)."""
        n, T = int(self.n), self.csg["T"]
        host = {k: v for k, v in self.host.items() if k != "touch"}  # derived from the boxes: load recomputes it
        fdt = np.float64 if self.precision == "float64" else np.float32
        for k in ("V0", "V1", "V2", "N"):  # no meshes: one placeholder row, as the reference writes for none
            host.setdefault(k, np.zeros((0, 3), fdt))
        for k in ("cstart", "cn", "cmesh", "tchunk"):
            host.setdefault(k, np.zeros(1, np.int32))
        ub = np.asarray(host["ub"], np.int64)
        fm = 1 + int(np.max(np.asarray(T["np_"])))
        ch0 = n + 1  # the first mesh-chunk item: after every navigator id (none here)
        su = ch0
        srf = (su, su + int(ub[-1]) * fm * 4, fm)
        parent = np.asarray(host["parent"][:n], np.int64)
        kids = tuple(tuple(int(c) for c in np.flatnonzero(parent == (v if v < n else n))) for v in range(n + 1))
        csg = _Record("lumen.geometry.csg_build:CsgTables", dict(
            T=dict(T), K=dict(self.csg["K"]), G=dict(self.csg["G"]), kind_box=(),
            empty_leaf=int(self.csg["empty_leaf"]), has_torus=bool(np.any(np.asarray(T["typ"]) == 2))))
        return _Record("lumen.geometry.navigator_build:NavArrays", dict(
            host=host, csg=csg, n=n, n_mesh=int(self.n_mesh), chunk=MESH_CHUNK, ch0=ch0, srf=srf,
            names=tuple(self.names), defs=(), stats=dict(self.stats), owned=(), clipped=(), kids=kids,
            lo_w=np.asarray(self.lo_w, np.float64), hi_w=np.asarray(self.hi_w, np.float64), probe={},
            origins=tuple(self.names) + ("OUT",), pieces=tuple(int(x) for x in np.asarray(host["piece"])[:n]),
            thin=(), precision=self.precision))

    @staticmethod
    def load(path: str) -> "NavArrays":
        """Read a cache entry: the reference's layout (``save``; an entry the reference's ``geometry/cache.py``
        ``_save`` wrote, whose arrays hold no meshes), or this build's earlier one (``host/<k>``, ``csg/<T|K|G>/<k>``
        arrays).  Raises ``ValueError`` for anything else."""  # noqa: D205, D209
        with open(os.path.join(path, "meta.json")) as f:
            meta = json.load(f)
        with np.load(os.path.join(path, "arrays.npz"), allow_pickle=False) as z:
            store = {k: z[k] for k in z.files}
        if isinstance(meta, dict) and "dc" in meta:
            rec = _value_of(meta, store)
            if not (isinstance(rec, _Record) and rec.cls.endswith(":NavArrays")):
                raise ValueError(f"navigator cache entry {path}: not a NavArrays record")
            f = rec.fields
            if int(f.get("n_mesh", 0)) > 0:
                raise ValueError(f"navigator cache entry {path}: holds {f['n_mesh']} meshes; this navigator has none")
            c = f["csg"].fields if isinstance(f["csg"], _Record) else f["csg"]
            csg = {"T": dict(c["T"]), "K": dict(c["K"]), "G": dict(c["G"]), "empty_leaf": int(c["empty_leaf"])}
            host = {k: v for k, v in f["host"].items() if k not in ("V0", "V1", "V2", "N", "cstart", "cn", "cmesh",
                                                                        "tchunk")}
            lo_w, hi_w = np.asarray(f["lo_w"]), np.asarray(f["hi_w"])
            if "touch" not in host:  # derived from the boxes: an entry written elsewhere lacks it
                host["touch"] = touch_flags(host["parent"], lo_w, hi_w)
            return NavArrays(host, csg, int(f["n"]), tuple(f["names"]), lo_w, hi_w,
                             0, str(f.get("precision", "float32")), dict(f.get("stats", {})))
        if isinstance(meta, dict) and meta.get("format") == "lumen-navarrays/1":
            host = {k[5:]: v for k, v in store.items() if k.startswith("host/")}
            csg = {tab: {k[len(tab) + 5:]: v for k, v in store.items() if k.startswith(f"csg/{tab}/")}
                   for tab in ("T", "K", "G")}
            csg["empty_leaf"] = meta.get("empty_leaf", len(csg["T"]["typ"]) - 1)
            if "touch" not in host:
                host["touch"] = touch_flags(host["parent"], store["lo_w"], store["hi_w"])
            return NavArrays(host, csg, int(meta["n"]), tuple(meta["names"]), store["lo_w"], store["hi_w"],
                             int(meta.get("n_mesh", 0)), meta.get("precision", "float32"), meta.get("stats", {}))
        raise ValueError(f"navigator cache entry {path}: an unknown layout")


FORMAT = "reference-encode/1"  # the entries' layout: the reference's _encode tree (a changed layout changes keys)
MESH_CHUNK = 8  # triangles per mesh chunk in the reference's format (no meshes here; the field is written as is)


@dataclass
class _Record:
    """A dataclass of the reference's format as read or written here: its ``module:qualname`` and fields."""

    cls: str
    fields: dict


def _tree_of(x, store: dict):
    """Return the JSON tree of a value in the reference's cache encoding: an array (or numpy scalar) stored under the
    next name ``a<i>`` as {"a": name, "scalar": bool}; a record as {"dc": class, "f": {field: tree}}; a dict as
    {"m": [[key tree, value tree], ...]}; a tuple {"t": [...]}, a list {"l": [...]}; None, bool, int, float and str
    as themselves."""  # noqa: D205, D209
    if isinstance(x, (np.ndarray, np.generic)):
        name = f"a{len(store)}"
        store[name] = np.asarray(x)
        return {"a": name, "scalar": isinstance(x, np.generic)}
    if isinstance(x, _Record):
        return {"dc": x.cls, "f": {k: _tree_of(v, store) for k, v in x.fields.items()}}
    if isinstance(x, dict):
        return {"m": [[_tree_of(k, store), _tree_of(v, store)] for k, v in x.items()]}
    if isinstance(x, tuple):
        return {"t": [_tree_of(v, store) for v in x]}
    if isinstance(x, list):
        return {"l": [_tree_of(v, store) for v in x]}
    if x is None or isinstance(x, (bool, int, float, str)):
        return x
    raise TypeError(f"navigator cache: cannot store a {type(x).__name__}")


def _value_of(t, store: dict):
    """Return the value of a tree ``_tree_of`` (or the reference's ``_encode``) wrote; a record as a ``_Record``."""
    if not isinstance(t, dict):
        return t
    if "a" in t:
        a = store[t["a"]]
        return a[()] if t.get("scalar") else a
    if "dc" in t:
        return _Record(t["dc"], {k: _value_of(v, store) for k, v in t["f"].items()})
    if "m" in t:
        return {_value_of(k, store): _value_of(v, store) for k, v in t["m"]}
    if "t" in t:
        return tuple(_value_of(v, store) for v in t["t"])
    return [_value_of(v, store) for v in t.get("l", [])]


# ---------------------------------------------------------------------------------------------- the CSG tables
def csg_tables(kinds: list[Solid], dtype=np.float32) -> dict:
    """Return the CSG tables of ``kinds``: T (leaves, one empty leaf past the end), K (leaf and program ranges), G.

    Columns as the reference's (``geometry/csg_build.py:140-165``): ``typ``, ``p0``, ``np_``, ``sa``, ``c``,
    ``R``, ``r``, ``o*``, ``b*``, ``q*``, ``bl*``, ``bh*``, ``A**``, ``m**`` per leaf; ``n*``, ``h`` per plane.
    """
    rows, planes, prog, krows = [], [], [], []
    for kd in kinds:
        if program_depth(kd.program) > MAX_DEPTH:
            raise ValueError(f"csg tables: solid {kd.name!r} is deeper than {MAX_DEPTH} operands")
        l0, p0 = len(rows), len(prog)
        for lf in kd.leaves:
            rows.append(dict(typ=lf.typ, o=lf.origin, q=lf.origin + lf.qo, A=lf.A, b=lf.b, c=lf.c,
                             sa=max(float(np.max(np.abs(lf.A))), 1e-30), p0=len(planes), np_=len(lf.N)))
            planes += [(lf.N[k], lf.H[k]) for k in range(len(lf.N))]
        prog += list(kd.program)
        krows.append((l0, len(kd.leaves), p0, len(kd.program)))
    empty_leaf = len(rows)
    rows.append(dict(typ=0, o=np.zeros(3), q=np.zeros(3), A=np.zeros((3, 3)), b=np.zeros(3), c=0.0, sa=1.0,
                     p0=len(planes), np_=0))
    f = lambda v: np.asarray(v, np.float64).astype(dtype)  # noqa: E731
    T = dict(typ=np.asarray([r["typ"] for r in rows], np.int32), p0=np.asarray([r["p0"] for r in rows], np.int32),
             np_=np.asarray([r["np_"] for r in rows], np.int32), sa=f([r["sa"] for r in rows]),
             c=f([r["c"] for r in rows]), R=f([r["A"][0, 0] if r["typ"] == 2 else 0.0 for r in rows]),
             r=f([r["A"][0, 1] if r["typ"] == 2 else 0.0 for r in rows]))
    for a in range(3):
        T[f"o{a}"] = f([r["o"][a] for r in rows])
        T[f"b{a}"] = f([r["b"][a] for r in rows])
        T[f"q{a}"] = f([r["q"][a] for r in rows])
        T[f"bl{a}"] = f(np.full(len(rows), -HUGE))
        T[f"bh{a}"] = f(np.full(len(rows), HUGE))
        for b in range(3):
            T[f"A{a}{b}"] = f([r["A"][a, b] for r in rows])
            T[f"m{a}{b}"] = f([1.0 if a == b else 0.0 for _ in rows])
    planes = planes or [(np.zeros(3), 0.0)]
    for a in range(3):
        T[f"n{a}"] = f([p[0][a] for p in planes])
    T["h"] = f([p[1] for p in planes])
    K = {k: np.asarray([x[j] for x in krows], np.int32) for j, k in enumerate(("l0", "nl", "p0", "pl"))}
    prog = prog or [(0, 0)]
    G = dict(op=np.asarray([g[0] for g in prog], np.int32), arg=np.asarray([g[1] for g in prog], np.int32))
    return dict(T=T, K=K, G=G, empty_leaf=empty_leaf)


# ---------------------------------------------------------------------------------------------- the skin core
SKIN_MARGIN_REL = 1e-3  # the core's clearance from the deepest daughter box, relative to the solid's half-size
SKIN_MAX_REL = 0.5  # no core when the daughters reach deeper than this share of the solid's smallest half-size


def _dist_to_hull_2d(pts):
    """Return the distance from the origin to the convex hull of 2D points (0 when inside it)."""
    from scipy.spatial import ConvexHull

    try:
        hull = ConvexHull(pts)
    except Exception:  # degenerate (collinear) points: the distance to their segments
        hull = None
    if hull is not None and np.all(hull.equations[:, 2] <= 0.0):  # the origin satisfies every facet: inside
        return 0.0
    best = np.inf
    for i in range(len(pts)):
        for j in range(i + 1, len(pts)):
            a, b = pts[i], pts[j]
            ab = b - a
            den = float(ab @ ab)
            s = 0.0 if den == 0.0 else min(max(-float(a @ ab) / den, 0.0), 1.0)
            best = min(best, float(np.linalg.norm(a + s * ab)))
    return best


def skin_core(kd: Solid, boxes) -> Solid | None:
    """Return the daughter-free core of solid ``kd`` (in its own frame) for daughters given as oriented boxes
    (centre C, axes U (columns), half-sizes h) in that frame, or None.

    ``kd`` must be one convex leaf: half-spaces n.(x - o) <= h, with or without a quadric that is a ball or a round
    cylinder about its centre (A's eigenvalues 0 or 1, b = 0, c = -r^2).  A box B lies outside the core shrunk by D
    when it lies beyond one shrunk face: a plane with min_B n.(x - o) >= h - D (the box's support), or the quadric
    with the least distance of B from the ball's centre (the box point nearest it) or from the cylinder's axis (the
    box projected along the axis: the distance to that convex polygon) at least r - D.  D is the largest over the
    daughters of the least such D over the faces, plus the margin.
    """
    if len(kd.leaves) != 1 or kd.program != [(0, 0)] or len(boxes) == 0:
        return None
    lf = kd.leaves[0]
    if lf.typ not in (0, 1):
        return None
    half = 0.5 * float(np.min(kd.hi - kd.lo))
    axes = None
    if lf.typ == 1:
        w, V = np.linalg.eigh(lf.A)
        unit, zero = np.abs(w - 1.0) < 1e-9, np.abs(w) < 1e-9
        if not (np.all(unit | zero) and unit.sum() >= 2 and np.allclose(lf.b, 0.0) and lf.c < 0.0):
            return None
        r, q, axes = float(np.sqrt(-lf.c)), lf.origin + lf.qo, V[:, unit]
    D = 0.0
    for C, U, h in boxes:
        need = np.inf
        for k in range(len(lf.N)):  # the box's least n.(x - o): its centre's less the support
            s = float(lf.N[k] @ (C - lf.origin)) - float(np.abs(lf.N[k] @ U) @ h)
            need = min(need, lf.H[k] - s)
        if axes is not None:
            if axes.shape[1] == 3:  # a ball: the box point nearest its centre
                s = U.T @ (q - C)
                dist = float(np.linalg.norm(s - np.clip(s, -h, h)))
            else:  # a cylinder: the box's corners projected on the plane across the axis
                corners = np.array([C + U @ (h * np.array([sx, sy, sz]))
                                    for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
                dist = _dist_to_hull_2d((corners - q) @ axes)
            need = min(need, r - dist)
        D = max(D, need)
    D += SKIN_MARGIN_REL * half
    if not (0.0 < D < SKIN_MAX_REL * half):
        return None
    core = Leaf(lf.typ, lf.origin.copy(), lf.qo.copy(), lf.A.copy(), lf.b.copy(),
                -(np.sqrt(-lf.c) - D) ** 2 if lf.typ == 1 else 0.0, lf.N.copy(), lf.H - D)
    return Solid([core], [(0, 0)], kd.lo + D, kd.hi - D, kd.name + "/core")


# ---------------------------------------------------------------------------------------------- the builder
def _world_box(kd: Solid, R, t):
    c, h = 0.5 * (kd.lo + kd.hi), 0.5 * (kd.hi - kd.lo)
    cw, hw = R @ c + t, np.abs(R) @ h
    pad = BOX_PAD_REL * (np.abs(cw) + hw) + BOX_PAD_ABS
    return cw - hw - pad, cw + hw + pad


MIN_GROUP = 8  # a rotation shared by at least this many of a mother's daughters (or by all) may be a frame group
GROUP_GAIN = 4.0  # ... when their world boxes are at least this times their own (summed surface terms)


def _surface_terms(h):
    return h[:, 0] * h[:, 1] + h[:, 1] * h[:, 2] + h[:, 2] * h[:, 0]


def frame_groups(kids: list[int], volumes: list[Volume], kinds: list[Solid]) -> list:
    """Return a mother's daughters as frame groups [(Rg, [daughters])], Rg the group frame's rotation (x_w = Rg x_g),
    None for the world frame (the reference's ``geometry/bvh.py:217-255`` ``frame_groups_all``): a rotation (equal
    entries) shared by at least MIN_GROUP daughters, or by all, that is not the identity and whose daughters' world
    boxes are at least GROUP_GAIN times their own boxes (summed surface terms: what a BVH pays for) is a group of
    its own, in the rotations' lexicographic order; the rest one world-frame group, last.  Daughters keep their
    order within a group."""  # noqa: D205, D209
    rows = np.stack([np.asarray(volumes[k].R, np.float64).reshape(9) for k in kids])
    uniq, inv, cnt = np.unique(rows, axis=0, return_inverse=True, return_counts=True)
    inv = inv.ravel()
    h = np.stack([0.5 * (kinds[volumes[k].kind].hi - kinds[volumes[k].kind].lo) for k in kids])
    hw = np.einsum("nij,nj->ni", np.abs(rows.reshape(-1, 3, 3)), h)
    tw, tb = _surface_terms(hw), _surface_terms(h)
    out, rest = [], np.ones(len(kids), bool)
    for g in range(len(uniq)):
        m = np.flatnonzero(inv == g)
        Rg = uniq[g].reshape(3, 3)
        if not (cnt[g] >= MIN_GROUP or cnt[g] == len(kids)) or np.allclose(Rg, np.eye(3), atol=1e-9):
            continue
        if np.sum(tw[m]) < GROUP_GAIN * np.sum(tb[m]):
            continue
        out.append((Rg, [kids[i] for i in m]))
        rest[m] = False
    if rest.any():
        out.append((None, [kids[i] for i in np.flatnonzero(rest)]))
    return out


def _group_box(kd: Solid, Rg, R, t):
    """Return a placed solid's box (lo, hi) in a group frame (x_g = Rg^T x_w), padded as the world boxes."""
    c, h = 0.5 * (kd.lo + kd.hi), 0.5 * (kd.hi - kd.lo)
    M = Rg.T @ R
    cg, hg = M @ c + Rg.T @ t, np.abs(M) @ h
    pad = BOX_PAD_REL * (np.abs(cg) + hg) + BOX_PAD_ABS
    return cg - hg - pad, cg + hg + pad


def build(volumes: list[Volume], kinds: list[Solid], outside_cls: int, precision="float32",
          leaf_size: int = bvh.LEAF_SIZE, framed: bool = True) -> NavArrays:
    """Return the NavArrays of placed ``volumes`` (in depth-first order: a mother before its daughters).
    ``framed``: rotated daughters in frame groups of their own (``frame_groups``), else every group in the world
    frame (the same navigator, looser boxes)."""  # noqa: D205, D209
    fdt = np.float32 if precision == "float32" else np.float64
    n = len(volumes)
    if n == 0:
        raise ValueError("navarrays.build: no volumes")
    parent = np.array([n if v.parent < 0 else v.parent for v in volumes], np.int64)
    for i, p in enumerate(parent):
        if p < n and p >= i:
            raise ValueError("navarrays.build: volumes must come after their mother (depth-first order)")
    lo_w, hi_w = np.zeros((n, 3)), np.zeros((n, 3))
    for i, v in enumerate(volumes):
        lo_w[i], hi_w[i] = _world_box(kinds[v.kind], v.R, v.t)
    kids = [[] for _ in range(n + 1)]
    for i in range(n):
        kids[parent[i]].append(i)
    # the daughter-free cores: extra kinds (no volume places them) whose one leaf each volume's "skin" names
    cores, core_of = [], np.full(n + 1, -1, np.int64)
    for u in range(n):
        if kids[u]:
            v = volumes[u]
            boxes = []
            for k in kids[u]:  # each daughter's own box, oriented, in the mother's frame
                kk, vk = kinds[volumes[k].kind], volumes[k]
                boxes.append((v.R.T @ (vk.R @ (0.5 * (kk.lo + kk.hi)) + vk.t - v.t), v.R.T @ vk.R,
                              0.5 * (kk.hi - kk.lo)))
            c = skin_core(kinds[v.kind], boxes)
            if c is not None:
                core_of[u] = len(kinds) + len(cores)
                cores.append(c)
    csg = csg_tables(list(kinds) + cores, fdt)
    nodes = dict(lo=[], hi=[], skip=[], start=[], count=[])
    items, g0, g1, vg0, vg1 = [], [], [], np.zeros(n + 1, np.int32), np.zeros(n + 1, np.int32)
    gR = []
    depth = 0
    for u in range(n + 1):
        vg0[u] = len(g0)
        groups = (frame_groups(kids[u], volumes, kinds) if framed else [(None, kids[u])]) if kids[u] else []
        for Rg, members in groups:
            ks = np.asarray(members, np.int64)
            if Rg is None:
                blo, bhi = lo_w[ks], hi_w[ks]
            else:
                bb = [_group_box(kinds[volumes[k].kind], Rg, volumes[k].R, volumes[k].t) for k in ks]
                blo, bhi = np.stack([b_[0] for b_ in bb]), np.stack([b_[1] for b_ in bb])
            tr = bvh.build(blo, bhi, leaf_size)
            base_n, base_i = len(nodes["lo"]), len(items)
            g0.append(base_n)
            gR.append(np.eye(3) if Rg is None else Rg)
            nodes["lo"] += list(tr["lo"])
            nodes["hi"] += list(tr["hi"])
            nodes["skip"] += list(tr["skip"] + base_n)
            nodes["start"] += list(np.where(tr["count"] > 0, tr["start"] + base_i, 0))
            nodes["count"] += list(tr["count"])
            items += list(ks[tr["items"]])
            g1.append(len(nodes["lo"]))
            depth = max(depth, tr["depth"])
        vg1[u] = len(g0)
    if depth > STACK:
        raise ValueError(f"navarrays.build: a BVH of depth {depth} exceeds the traversal stack ({STACK})")
    G = len(g0)
    size = np.ones(n, np.int64)
    for i in range(n - 1, -1, -1):
        if parent[i] < n:
            size[parent[i]] += size[i]
    nl = np.maximum(csg["K"]["nl"][[v.kind for v in volumes]], 1)
    host = dict(
        lo=np.asarray(nodes["lo"] or np.zeros((1, 3))).astype(fdt),
        hi=np.asarray(nodes["hi"] or np.zeros((1, 3))).astype(fdt),
        skip=np.asarray(nodes["skip"] or [0], np.int32), start=np.asarray(nodes["start"] or [0], np.int32),
        count=np.asarray(nodes["count"] or [0], np.int32), items=np.asarray(items or [0], np.int32),
        gR=(np.stack(gR) if gR else np.eye(3)[None]).astype(fdt),
        g0=np.asarray(g0 or [0], np.int32), g1=np.asarray(g1 or [0], np.int32),
        vg0=vg0, vg1=vg1, vg1u=vg1.copy(),
        parent=np.concatenate([parent, [n]]).astype(np.int32),
        cls=np.asarray([v.cls for v in volumes] + [outside_cls], np.int32),
        kind=np.asarray([v.kind for v in volumes], np.int32),
        R=np.stack([v.R for v in volumes]).astype(fdt), t=np.stack([v.t for v in volumes]).astype(fdt),
        prio=np.arange(n + 1, dtype=np.int32),
        size=np.concatenate([size, [0]]).astype(np.int32),
        piece=np.full(n + 1, -1, np.int32),
        ub=np.concatenate([[0], np.cumsum(nl)]).astype(np.int32),
        skin=np.where(core_of >= 0, csg["K"]["l0"][np.maximum(core_of, 0)], -1).astype(np.int32),
    )
    host["touch"] = touch_flags(parent, lo_w, hi_w)
    framed_groups = int(sum(not np.array_equal(r, np.eye(3)) for r in gR))
    stats = dict(volumes=n, kinds=len(kinds), groups=G, framed=framed_groups, bvh_depth=depth,
                 max_daughters=max(len(k) for k in kids), skin_cores=len(cores))
    return NavArrays(host, csg, n, tuple(v.name for v in volumes), lo_w, hi_w, 0, precision, stats)


def touch_flags(parent, lo_w, hi_w) -> np.ndarray:
    """Return (n+1,) int32: 1 where a sibling's world box lies within the relocation step of the volume's own, K_ON
    float32 bands at the geometry's coordinate scale (the largest coordinate plus the diagonal: the scale of any
    crossing): an exit from such a volume may land in that sibling.  Boxes hold their solids, so a volume whose box
    no sibling's box comes near can meet no sibling past its faces."""  # noqa: D205, D209
    n = len(lo_w)
    out = np.zeros(n + 1, np.int32)
    if n == 0:
        return out
    scale = float(np.max(np.abs(np.concatenate([lo_w, hi_w])))) + float(np.linalg.norm(hi_w.max(0) - lo_w.min(0)))
    tol = 4.0 * (8.0 * 2.0**-23 * scale + 1e-9)
    par = np.asarray(parent)[:n]
    for p in np.unique(par):
        ks = np.flatnonzero(par == p)
        if len(ks) < 2:
            continue
        lo, hi = lo_w[ks] - tol, hi_w[ks] + tol
        for a in range(0, len(ks), 512):  # pairwise box overlaps, in slabs
            b = slice(a, a + 512)
            ov = np.all((lo[b, None, :] <= hi[None, :, :]) & (lo[None, :, :] <= hi[b, None, :]), axis=2)
            ov[np.arange(ov.shape[0]), np.arange(a, a + ov.shape[0])] = False  # not itself
            out[ks[b]] |= ov.any(axis=1).astype(np.int32)
    return out


# ---------------------------------------------------------------------------------------------- the build cache
KIND = "nav"  # an entry's directory: <cache root>/build/nav-<key>
FORMAT = "lumen-navarrays/1"
_BUILDER_CODE = ("navarrays.py", "bvh.py", os.path.join("..", "csrc", "bvh_build.cpp"))  # what builds the arrays


def cache_root() -> str:
    """Return the per-user cache root (the reference's ``_env.cache_dir``): ``LUMEN_CACHE``, else ``lumen-<user>`` in
    the system's temporary directory."""  # noqa: D205, D209
    value = os.environ.get("LUMEN_CACHE")
    return value if value else os.path.join(tempfile.gettempdir(), f"lumen-{getpass.getuser()}")


def _digest_inputs(h, volumes, kinds, outside_cls, options):
    """Feed a canonical encoding of the builder's inputs into hash h (arrays by dtype, shape and bytes)."""
    def arr(x, dt=np.float64):
        a = np.ascontiguousarray(np.asarray(x, dt))
        h.update(f"{a.dtype.str}{a.shape}".encode())
        h.update(a.tobytes())

    h.update(f"{FORMAT}|{len(volumes)}|{len(kinds)}|{int(outside_cls)}|".encode())
    h.update(json.dumps(options, sort_keys=True).encode())
    for v in volumes:
        h.update(f"|v|{v.name}|{int(v.kind)}|{int(v.parent)}|{int(v.cls)}|".encode())
        arr(v.R)
        arr(v.t)
    for kd in kinds:
        h.update(f"|k|{len(kd.leaves)}|".encode())
        arr(np.asarray(kd.program, np.int64).reshape(-1, 2), np.int64)
        arr(kd.lo)
        arr(kd.hi)
        for lf in kd.leaves:
            h.update(f"|l|{int(lf.typ)}|".encode())
            for x in (lf.origin, lf.qo, lf.A, lf.b, [lf.c], np.asarray(lf.N).reshape(-1, 3), lf.H):
                arr(x)


def key(volumes: list[Volume], kinds: list[Solid], outside_cls: int, **options) -> str:
    """Return the cache key of the arrays ``build`` makes of these inputs (the reference's ``geometry/cache.py:175-183``
    ``key``): sha256 of the inputs, the options and the builder's own source files."""  # noqa: D205, D209
    h = hashlib.sha256()
    _digest_inputs(h, volumes, kinds, outside_cls, options)
    here = os.path.dirname(os.path.abspath(__file__))
    for f in _BUILDER_CODE:  # a copy without a source file keys by the built library's bytes instead (below)
        path = os.path.join(here, f)
        if os.path.isfile(path):
            with open(path, "rb") as src:
                h.update(hashlib.sha256(src.read()).digest())
        else:
            h.update(f"missing:{f}".encode())
    lib = os.path.join(os.path.dirname(here), "liblumen_hip.so")  # the native BVH builder (lh_bvh_build)
    if os.path.isfile(lib):
        with open(lib, "rb") as so:
            h.update(hashlib.sha256(so.read()).digest())
    h.update(FORMAT.encode())
    h.update(np.__version__.encode())
    return h.hexdigest()


def cached(volumes: list[Volume], kinds: list[Solid], outside_cls: int, precision="float32",
           leaf_size: int = bvh.LEAF_SIZE, root: str | None = None) -> tuple[NavArrays, str]:
    """Return (the NavArrays of ``build``, their cache key): read from ``<root>/build/nav-<key>`` (``save``'s
    ``arrays.npz`` + ``meta.json``), else built and stored there (the reference's ``geometry/cache.py:238-266``
    ``nav_arrays``).  An entry is written into a temporary directory beside it and renamed into place, so a reader
    never sees a partial one and concurrent builders of one key leave one entry."""  # noqa: D205, D209
    options = dict(precision=precision, leaf_size=int(leaf_size), framed=True)
    k = key(volumes, kinds, outside_cls, **options)
    base = os.path.join(root or cache_root(), "build")
    path = os.path.join(base, f"{KIND}-{k}")
    if os.path.isfile(os.path.join(path, "meta.json")):
        try:
            return NavArrays.load(path), k
        except (OSError, ValueError, KeyError, TypeError):  # a damaged or foreign entry: built again, replaced
            shutil.rmtree(path, ignore_errors=True)
    nav = build(volumes, kinds, outside_cls, precision, leaf_size)
    os.makedirs(base, exist_ok=True)
    tmp = tempfile.mkdtemp(prefix=f".{KIND}-", dir=base)
    try:
        nav.save(tmp)
        try:
            os.rename(tmp, path)
        except OSError:  # another builder stored the same key first
            pass
    finally:
        if os.path.isdir(tmp):
            shutil.rmtree(tmp, ignore_errors=True)
    return nav, k
