"""Solids as CSG leaves: a quadric intersected with half-spaces, combined by a postfix boolean program.

A leaf is ``{x : (x - q)'A (x - q) + 2 b.(x - q) + c <= 0} & {n_k . (x - o) <= h_k}`` in its solid's frame, with
``q = o + qo`` (the quadric's own centre, so a small quadric far from the solid's origin keeps coefficients of its
own size).  Along a line each leaf is at most two intervals.  A solid ("kind") is its leaves plus a postfix program
(op 0 pushes leaf ``arg``; 1 union, 2 intersection, 3 subtraction), the same leaf/program model the reference's
navigator tables hold (``geometry/csg_build.py:1-46``, ``geometry/solids.py:1-40``).

Constructors take lengths in metres.  Every constructor attaches an exact axis-aligned box; booleans combine boxes
by the CSG rules (union: hull, intersection: overlap, subtraction: the left operand's).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

OPS = {"union": 1, "inter": 2, "sub": 3}


@dataclass
class Leaf:
    """One CSG leaf (see the module docstring); ``typ`` 0 planes only, 1 quadric and planes."""

    typ: int
    origin: np.ndarray
    qo: np.ndarray
    A: np.ndarray
    b: np.ndarray
    c: float
    N: np.ndarray  # (k, 3) outward unit normals
    H: np.ndarray  # (k,)

    def moved(self, R: np.ndarray, t: np.ndarray) -> "Leaf":
        """Return the leaf placed by x' = R x + t (R a rotation)."""
        return Leaf(self.typ, R @ self.origin + t, R @ self.qo, R @ self.A @ R.T, R @ self.b, self.c,
                    self.N @ R.T if len(self.N) else self.N.copy(), self.H.copy())

    def inside(self, x: np.ndarray) -> np.ndarray:
        """Return float64 membership of points x (..., 3)."""
        x = np.asarray(x, np.float64)
        ok = np.ones(x.shape[:-1], bool)
        if self.typ == 1:
            y = x - (self.origin + self.qo)
            q = np.einsum("...i,ij,...j->...", y, self.A, y) + 2.0 * y @ self.b + self.c
            ok &= q <= 0.0
        if len(self.N):
            ok &= np.all((x - self.origin) @ self.N.T - self.H <= 0.0, -1)
        return ok


@dataclass
class Solid:
    """A CSG solid: leaves, a postfix program over them, and its box (lo, hi) in its own frame."""

    leaves: list
    program: list  # [(op, arg)]
    lo: np.ndarray
    hi: np.ndarray
    name: str = ""
    params: dict = field(default_factory=dict)

    def inside(self, x: np.ndarray) -> np.ndarray:
        """Return float64 membership of points x (..., 3) by the postfix program."""
        vals = [lf.inside(x) for lf in self.leaves]
        st = []
        for op, arg in self.program:
            if op == 0:
                st.append(vals[arg])
            else:
                b, a = st.pop(), st.pop()
                st.append(a | b if op == 1 else a & b if op == 2 else a & ~b)
        return st[-1]


def _z3():
    return np.zeros(3)


def _planes_leaf(N, H, origin=None) -> Leaf:
    N = np.asarray(N, np.float64).reshape(-1, 3)
    return Leaf(0, _z3() if origin is None else np.asarray(origin, float), _z3(), np.zeros((3, 3)), _z3(), 0.0,
                N / np.linalg.norm(N, axis=1, keepdims=True), np.asarray(H, np.float64).reshape(-1))


def box(hx: float, hy: float, hz: float, name="box") -> Solid:
    """Return a box of half-lengths (hx, hy, hz) centred at the origin."""
    N = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], float)
    H = np.array([hx, hx, hy, hy, hz, hz], float)
    h = np.array([hx, hy, hz], float)
    return Solid([_planes_leaf(N, H)], [(0, 0)], -h, h, name, dict(hx=hx, hy=hy, hz=hz))


def cylinder_leaf(r: float, hz: float) -> Leaf:
    """Return the full cylinder x^2 + y^2 <= r^2, |z| <= hz as one leaf."""
    return Leaf(1, _z3(), _z3(), np.diag([1.0, 1.0, 0.0]), _z3(), -r * r,
                np.array([[0, 0, 1.0], [0, 0, -1.0]]), np.array([hz, hz]))


def tube(rmax: float, hz: float, rmin: float = 0.0, name="tube") -> Solid:
    """Return a full-phi tube (G4Tubs without a phi cut): outer radius rmax, half length hz, inner radius rmin."""
    lo, hi = np.array([-rmax, -rmax, -hz]), np.array([rmax, rmax, hz])
    if rmin <= 0.0:
        return Solid([cylinder_leaf(rmax, hz)], [(0, 0)], lo, hi, name, dict(rmax=rmax, hz=hz))
    inner = Leaf(1, _z3(), _z3(), np.diag([1.0, 1.0, 0.0]), _z3(), -rmin * rmin, np.zeros((0, 3)), np.zeros(0))
    return Solid([cylinder_leaf(rmax, hz), inner], [(0, 0), (0, 1), (3, 0)], lo, hi, name,
                 dict(rmax=rmax, hz=hz, rmin=rmin))


def orb(r: float, name="orb") -> Solid:
    """Return a full sphere of radius r."""
    lf = Leaf(1, _z3(), _z3(), np.eye(3), _z3(), -r * r, np.zeros((0, 3)), np.zeros(0))
    h = np.full(3, r)
    return Solid([lf], [(0, 0)], -h, h, name, dict(r=r))


def sphere_cap(r: float, height: float, name="cap") -> Solid:
    """Return the cap of a sphere of radius r (centred at the origin) above the plane z = r - height: one leaf.

    The cap's flat face is at z = r - height and its dome reaches z = r (the cap points along +z).
    """
    z0 = r - height
    lf = Leaf(1, _z3(), _z3(), np.eye(3), _z3(), -r * r, np.array([[0, 0, -1.0]]), np.array([-z0]))
    rc = np.sqrt(max(r * r - z0 * z0, 0.0)) if z0 > 0 else r
    return Solid([lf], [(0, 0)], np.array([-rc, -rc, z0]), np.array([rc, rc, r]), name, dict(r=r, height=height))


def polar_cone(theta: float, r: float, name="cone") -> Solid:
    """Return the points whose polar angle (from +z, at the origin) is at most ``theta`` (0 < theta < pi), boxed by
    +-r.  Below pi/2 one leaf: the cone (x^2 + y^2) cos^2 - z^2 sin^2 <= 0 cut by z >= 0; at pi/2 the half-space
    z >= 0; above it the union of that half-space and the outside of the cone of half-angle pi - theta."""  # noqa: D205, D209
    h = np.full(3, float(r))
    up = np.array([[0.0, 0.0, -1.0]])
    half = _planes_leaf(up, [0.0])
    if abs(theta - 0.5 * np.pi) < 1e-15:
        return Solid([half], [(0, 0)], -h, h, name, dict(theta=theta))
    if theta < 0.5 * np.pi:
        c2, s2 = np.cos(theta) ** 2, np.sin(theta) ** 2
        lf = Leaf(1, _z3(), _z3(), np.diag([c2, c2, -s2]), _z3(), 0.0, up, np.array([0.0]))
        return Solid([lf], [(0, 0)], -h, h, name, dict(theta=theta))
    a = np.pi - theta
    c2, s2 = np.cos(a) ** 2, np.sin(a) ** 2
    out = Leaf(1, _z3(), _z3(), np.diag([-c2, -c2, s2]), _z3(), 0.0, np.zeros((0, 3)), np.zeros(0))
    return Solid([half, out], [(0, 0), (0, 1), (OPS["union"], 0)], -h, h, name, dict(theta=theta))


def sphere(rmin: float, rmax: float, theta0: float = 0.0, dtheta: float = np.pi, name="sphere") -> Solid:
    """Return G4Sphere without a phi cut: radii [rmin, rmax], polar angles [theta0, theta0 + dtheta] (radians).

    The shell (one orb minus another), intersected with the cone of the upper polar angle and less the cone of the
    lower one where they cut; its box is exact: z over the corners (r, theta), the radius in xy at most rmax times
    the largest sin theta of the range.
    """
    t1 = min(theta0 + dtheta, np.pi)
    if not (0.0 <= rmin < rmax and 0.0 <= theta0 < t1):
        raise ValueError(f"sphere {name!r}: 0 <= rmin < rmax and 0 <= theta0 < theta0 + dtheta, got "
                         f"{rmin}, {rmax}, {theta0}, {dtheta}")
    s = orb(rmax, name=name)
    if rmin > 0.0:
        s = boolean("sub", s, orb(rmin), name=name)
    if t1 < np.pi - 1e-12:
        s = boolean("inter", s, polar_cone(t1, rmax), name=name)
    if theta0 > 1e-12:
        s = boolean("sub", s, polar_cone(theta0, rmax), name=name)
    zs = [r * np.cos(t) for r in (rmin, rmax) for t in (theta0, t1)]
    rho = rmax * (1.0 if theta0 <= 0.5 * np.pi <= t1 else max(np.sin(theta0), np.sin(t1)))
    s.lo, s.hi = np.array([-rho, -rho, min(zs)]), np.array([rho, rho, max(zs)])
    s.params = dict(rmin=rmin, rmax=rmax, theta0=theta0, dtheta=dtheta)
    return s


def _box_corners(lo, hi):
    return np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])


def boolean(op: str, a: Solid, b: Solid, R=None, t=None, name="boolean") -> Solid:
    """Return ``a op b`` with b placed in a's frame by x_a = R x_b + t (GDML's boolean second-solid pose)."""
    R = np.eye(3) if R is None else np.asarray(R, float)
    t = _z3() if t is None else np.asarray(t, float)
    bl = [lf.moved(R, t) for lf in b.leaves]
    n = len(a.leaves)
    prog = list(a.program) + [(o, arg + n if o == 0 else arg) for o, arg in b.program] + [(OPS[op], 0)]
    cb = _box_corners(b.lo, b.hi) @ R.T + t
    blo, bhi = cb.min(0), cb.max(0)
    if op == "union":
        lo, hi = np.minimum(a.lo, blo), np.maximum(a.hi, bhi)
    elif op == "inter":
        lo, hi = np.maximum(a.lo, blo), np.minimum(a.hi, bhi)
    else:
        lo, hi = a.lo.copy(), a.hi.copy()
    return Solid(list(a.leaves) + bl, prog, lo, hi, name)


def program_depth(program) -> int:
    """Return the operand-stack depth a postfix program needs."""
    d = m = 0
    for op, _ in program:
        d = d + 1 if op == 0 else d - 1
        m = max(m, d)
    return m
