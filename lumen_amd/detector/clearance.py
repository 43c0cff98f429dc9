"""The diffusion jump's clearance: a proven lower bound, on a regular grid, of the distance to every surface.

Restates the rules of the reference's ``geometry/clearance.py`` (and the grid choices of ``transport/fastpath.py``
``prepare``) for this build's navigator arrays, in float64 on the host:

- a volume whose world box does not hold the node: the distance to that box (its solid lies inside it);
- else the least, over the volume's CSG leaves (a boolean's surface lies in the union of its leaves' surfaces), of
    plane n.(x - o) <= h:            |n.(x - o) - h| / |n|                                      (exact)
    quadric f = y'Ay + 2b.y + c:     the root r >= 0 of ||A|| r^2 + |grad f| r = |f|              (|f(y + e) - f(y)|
                                     <= |grad f| |e| + ||A|| |e|^2, so f keeps its sign within r)
    torus (R, r about axis b):       | sqrt((rho - R)^2 + z^2) - r |                               (exact)
- the least over every volume, at most ``cap``, less the navigator's band at the grid's coordinate scale, rounded
  down to float32.

The device reads c(x) = clear[node] - |x - node| at the node nearest x: again a lower bound (the distance is
1-Lipschitz), so any spacing is exact; the spacing changes only how often and how far a walk goes.  A medium is
eligible when its only channel is Rayleigh (no Mie, no wavelength shifter); its grid spacing is 2 R_0 lambda / sqrt(3)
with lambda its shortest Rayleigh length, over the boxes of its volumes at least R_0 lambda half-wide, coarsened while
the grid exceeds MAX_NODES.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

R_0 = 4.0  # scattering lengths: the smallest ball a grid is built for (speed only)
MAX_NODES = 1 << 24  # float32 nodes of one medium's grid (64 MiB): a resource limit, speed only
COARSEN = 1.0 + 2.0**-20  # the least factor the spacing grows by while the grid does not fit
BAND_ULPS, BAND_FLOOR, EPS32 = 8.0, 1e-9, float(np.finfo(np.float32).eps)  # the navigator's band (nav.h band_at)


@dataclass
class JumpGrid:
    """One medium's clearance grid: nodes lo + h i, i < n (per axis), the bound ``clear`` (float32, shape n)."""

    medium: int
    lo: np.ndarray
    h: float
    n: np.ndarray
    clear: np.ndarray
    shortest_rayleigh: float


def _leaf(T, li):
    pl = slice(int(T["p0"][li]), int(T["p0"][li] + T["np_"][li]))
    return dict(typ=int(T["typ"][li]), o=np.array([T[f"o{a}"][li] for a in range(3)], np.float64),
                q=np.array([T[f"q{a}"][li] for a in range(3)], np.float64),
                A=np.array([[T[f"A{a}{b}"][li] for b in range(3)] for a in range(3)], np.float64),
                b=np.array([T[f"b{a}"][li] for a in range(3)], np.float64), c=float(T["c"][li]),
                N=np.stack([T["n0"], T["n1"], T["n2"]], 1)[pl].astype(np.float64),
                H=np.asarray(T["h"][pl], np.float64))


def leaf_bound(lf, xl):
    """Return a lower bound (N,) of the distance from points xl (N, 3; the leaf's frame) to the leaf's surfaces."""
    out = np.full(len(xl), np.inf)
    if len(lf["N"]):
        d = np.abs((xl - lf["o"]) @ lf["N"].T - lf["H"]) / np.linalg.norm(lf["N"], axis=1)
        out = np.minimum(out, d.min(axis=1))
    if lf["typ"] == 1:
        y = xl - lf["q"]
        A = 0.5 * (lf["A"] + lf["A"].T)
        f = np.einsum("ni,ij,nj->n", y, A, y) + 2.0 * y @ lf["b"] + lf["c"]
        gr = np.linalg.norm(2.0 * (y @ A + lf["b"]), axis=1)
        an, af = float(np.max(np.abs(np.linalg.eigvalsh(A)))), np.abs(f)
        if an > 0:
            r = 2.0 * af / (gr + np.sqrt(gr * gr + 4.0 * an * af))
        else:
            r = np.where(gr > 0, af / np.maximum(gr, 1e-300), np.where(af > 0, np.inf, 0.0))
        out = np.minimum(out, r)
    elif lf["typ"] == 2:  # a torus: the distance to its surface, | sqrt((rho - R)^2 + z^2) - r |, exactly
        y = xl - lf["q"]
        z = y @ lf["b"]
        rho = np.sqrt(np.maximum((y * y).sum(1) - z * z, 0.0))
        out = np.minimum(out, np.abs(np.hypot(rho - lf["A"][0, 0], z) - lf["A"][0, 1]))
    return out


def grid(det, volumes, h):
    """Return (lo, n) of the grid of spacing h whose nodes cover the union of the volumes' world boxes."""
    lo = np.min(np.asarray(det.nav.lo_w, np.float64)[list(volumes)], 0)
    hi = np.max(np.asarray(det.nav.hi_w, np.float64)[list(volumes)], 0)
    return lo, np.floor((hi - lo) / h).astype(np.int64) + 2


def bound(det, lo, h, n, cap):
    """Return the proven clearance (float32, shape n) at the nodes lo + h i, at most ``cap`` before the band."""
    nav = det.nav
    T, K = nav.csg["T"], nav.csg["K"]
    axes = [lo[a] + h * np.arange(int(n[a])) for a in range(3)]
    X = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    out = np.full(len(X), float(cap))
    R, t, kind = np.asarray(nav.host["R"], np.float64), np.asarray(nav.host["t"], np.float64), nav.host["kind"]
    for v in range(nav.n):
        blo, bhi = np.asarray(nav.lo_w[v], np.float64), np.asarray(nav.hi_w[v], np.float64)
        o = np.maximum(np.maximum(blo - X, X - bhi), 0.0)
        dbox = np.sqrt((o * o).sum(1))
        out = np.minimum(out, np.where(dbox > 0, dbox, np.inf))
        inside = np.flatnonzero(dbox == 0.0)
        if not len(inside):
            continue
        xl = (X[inside] - t[v]) @ R[v]  # R^T (x - t): the volume's frame
        k = int(kind[v])
        lb = np.full(len(inside), np.inf)
        for li in range(int(K["l0"][k]), int(K["l0"][k] + K["nl"][k])):
            lb = np.minimum(lb, leaf_bound(_leaf(T, li), xl))
        out[inside] = np.minimum(out[inside], lb)
    scale = float(np.max(np.abs(np.concatenate([lo, lo + h * (np.asarray(n) - 1)])))) + cap
    out = out - (BAND_ULPS * EPS32 * scale + BAND_FLOOR)
    c32 = out.astype(np.float32)
    c32 = np.where(c32.astype(np.float64) > out, np.nextafter(c32, np.float32(-np.inf)), c32)  # rounded down
    return np.maximum(c32, np.float32(0.0)).reshape(tuple(int(x) for x in n))


def eligible(med) -> bool:
    """Whether a medium's only event channel is Rayleigh (the walk inside the ball is its plain walk)."""
    return med.rindex is not None and med.rayleigh is not None and med.mie_length is None and med.wls_length is None


def shortest_rayleigh(det, params, m) -> float:
    """Return medium m's shortest Rayleigh length (m) over its table, as packed (Geant4's water rule included)."""
    from .. import pack

    med = det.media[m]
    if med.rayleigh == "water_rule":
        rt = params[med.rindex]
        return float(np.min(pack.water_rayleigh(rt.energy, rt.value)))
    v = params[med.rayleigh]
    return float(np.min(np.asarray(getattr(v, "value", v), np.float64)))


def _volumes_by_medium(det):
    """Return {medium: (volume ids, their world boxes' half-widths along the narrowest axis)}."""
    nav = det.nav
    medium_of_volume = np.asarray(det.class_medium)[np.asarray(nav.host["cls"])[: nav.n]]
    extent = np.asarray(nav.hi_w, np.float64) - np.asarray(nav.lo_w, np.float64)
    narrow = 0.5 * extent.min(axis=1)
    return {int(m): (np.flatnonzero(medium_of_volume == m), narrow[medium_of_volume == m])
            for m in np.unique(medium_of_volume) if m >= 0}  # -1: a class without a medium


def _fitting_grid(det, volumes, spacing):
    """Return (lo, n, spacing) of a grid over the volumes' boxes at the given spacing or, when that holds more than
    MAX_NODES, at the least spacing (grown by at least COARSEN a try) whose grid does."""  # noqa: D205, D209
    lo, n = grid(det, volumes, spacing)
    while int(np.prod(n)) > MAX_NODES:
        spacing *= max((float(np.prod(n)) / MAX_NODES) ** (1.0 / 3.0), COARSEN)
        lo, n = grid(det, volumes, spacing)
    return lo, n, spacing


def jump_grids(det, params=None) -> list:
    """Return the JumpGrids of a detector's eligible media (the rules of the reference's ``fastpath.prepare``,
    ``transport/fastpath.py:117-152``).  A medium gets a grid when one of its volumes is wide enough to hold a
    ball of R_0 of its shortest Rayleigh lengths; the grid spans those volumes at spacing 2 R_0 lambda / sqrt(3)
    and its bound is capped at the widest of the medium's volumes; a grid whose largest clearance still holds no
    such ball is dropped."""  # noqa: D205, D209
    p = det.params() if params is None else params
    grids = []
    for m, (vols, narrow) in sorted(_volumes_by_medium(det).items()):
        if not eligible(det.media[m]):
            continue
        lam = shortest_rayleigh(det, p, m)
        ball = R_0 * lam
        wide = vols[narrow >= ball]
        if not (lam > 0 and len(wide)):
            continue
        lo, n, spacing = _fitting_grid(det, wide, 2.0 * ball / np.sqrt(3.0))
        clear = bound(det, lo, spacing, n, float(narrow.max()))
        if float(clear.max()) >= ball:
            grids.append(JumpGrid(m, lo, float(spacing), np.asarray(n, np.int64), clear, lam))
    return grids
