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
        A = self.A.copy() if self.typ == 2 else R @ self.A @ R.T  # a torus keeps its radii, its axis b turns
        return Leaf(self.typ, R @ self.origin + t, R @ self.qo, A, R @ self.b, self.c,
                    self.N @ R.T if len(self.N) else self.N.copy(), self.H.copy())

    def inside(self, x: np.ndarray) -> np.ndarray:
        """Return float64 membership of points x (..., 3)."""
        x = np.asarray(x, np.float64)
        ok = np.ones(x.shape[:-1], bool)
        if self.typ == 1:
            y = x - (self.origin + self.qo)
            q = np.einsum("...i,ij,...j->...", y, self.A, y) + 2.0 * y @ self.b + self.c
            ok &= q <= 0.0
        elif self.typ == 2:
            ok &= torus_g(x - (self.origin + self.qo), self.b, self.A[0, 0], self.A[0, 1]) <= 0.0
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


def torus_g(y, axis, R: float, r: float):
    """Return |y|^2 - 2 R rho + R^2 - r^2 (rho the distance from the axis): <= 0 inside the torus of swept radius
    R and tube radius r about ``axis`` through the origin; it has the sign of (rho - R)^2 + z^2 - r^2."""  # noqa: D205, D209
    y = np.asarray(y, np.float64)
    z = y @ np.asarray(axis, np.float64)
    rho = np.sqrt(np.maximum((y * y).sum(-1) - z * z, 0.0))
    return (y * y).sum(-1) - 2.0 * R * rho + R * R - r * r


def torus_leaf(R: float, r: float) -> Leaf:
    """A solid torus about z (swept radius R, tube radius r): a leaf of type 2, its radii in A[0, 0] and A[0, 1]
    (the tables' ``A00``, ``A01``, and their ``R``, ``r`` columns) and its axis in b."""  # noqa: D205, D209
    A = np.zeros((3, 3))
    A[0, 0], A[0, 1], A[1, 0] = R, r, r
    return Leaf(2, _z3(), _z3(), A, np.array([0.0, 0.0, 1.0]), 0.0, np.zeros((0, 3)), np.zeros(0))


def torus(rmin: float, rmax: float, rtor: float, phi0: float = 0.0, dphi: float = 2 * np.pi,
          name="torus") -> Solid:
    """Return G4Torus: the tube of radii [rmin, rmax] swept at radius rtor about z, azimuths [phi0, phi0 + dphi]."""
    if not (0.0 <= rmin < rmax and rtor > 0.0):
        raise ValueError(f"torus {name!r}: 0 <= rmin < rmax and rtor > 0")
    h = np.array([rtor + rmax, rtor + rmax, rmax])
    s = Solid([torus_leaf(rtor, rmax)], [(0, 0)], -h, h, name)
    if rmin > 0.0:
        s = boolean("sub", s, Solid([torus_leaf(rtor, rmin)], [(0, 0)], -h, h), name=name)
    if dphi < 2 * np.pi - 1e-12:
        s = boolean("inter", s, _phi_wedge(phi0, dphi, rtor + rmax, rmax), name=name)
    s.lo, s.hi = -h, h
    s.params = dict(rmin=rmin, rmax=rmax, rtor=rtor, phi0=phi0, dphi=dphi)
    return s


def _cone_leaf(r1: float, r2: float, hz: float) -> Leaf:
    """The full cone (frustum) x^2 + y^2 <= rho(z)^2, rho linear from r1 at z = -hz to r2 at z = +hz, |z| <= hz.

    rho = a + b z with a = (r1 + r2) / 2, b = (r2 - r1) / (2 hz): x^2 + y^2 - b^2 z^2 - 2 a b z - a^2 <= 0; rho >= 0
    on the slab (r1, r2 >= 0), so the quadric's other nappe lies outside it."""  # noqa: D205, D209
    a, b = 0.5 * (r1 + r2), (r2 - r1) / (2.0 * hz)
    return Leaf(1, _z3(), _z3(), np.diag([1.0, 1.0, -b * b]), np.array([0.0, 0.0, -a * b]), -a * a,
                np.array([[0, 0, 1.0], [0, 0, -1.0]]), np.array([hz, hz]))


def _phi_wedge(phi0: float, dphi: float, r: float, hz: float):
    """Return the solid of the points whose azimuth lies in [phi0, phi0 + dphi] (0 < dphi < 2 pi), boxed by
    (+-r, +-r, +-hz): one leaf of two half-planes through the z axis when dphi <= pi, else the union of two."""  # noqa: D205, D209
    h = np.array([r, r, hz])

    def half(phi):  # the points on the left of the direction phi: -sin x + cos y >= 0
        return np.array([np.sin(phi), -np.cos(phi), 0.0])

    def right(phi):  # the points on the right of the direction phi
        return -half(phi)

    if dphi <= np.pi:
        lf = _planes_leaf(np.stack([half(phi0), right(phi0 + dphi)]), [0.0, 0.0])
        return Solid([lf], [(0, 0)], -h, h, "wedge")
    a = _planes_leaf(half(phi0)[None], [0.0])
    b = _planes_leaf(right(phi0 + dphi)[None], [0.0])
    return Solid([a, b], [(0, 0), (0, 1), (OPS["union"], 0)], -h, h, "wedge")


def cons(rmin1: float, rmax1: float, rmin2: float, rmax2: float, hz: float, phi0: float = 0.0,
         dphi: float = 2 * np.pi, name="cons") -> Solid:
    """Return G4Cons: radii [rmin1, rmax1] at z = -hz and [rmin2, rmax2] at z = +hz, linear between, azimuths
    [phi0, phi0 + dphi] (a full turn: no cut)."""  # noqa: D205, D209
    if not (0.0 <= rmin1 <= rmax1 and 0.0 <= rmin2 <= rmax2 and rmax1 + rmax2 > 0.0 and hz > 0.0):
        raise ValueError(f"cons {name!r}: 0 <= rmin <= rmax at both ends and hz > 0")
    rmax = max(rmax1, rmax2)
    h = np.array([rmax, rmax, hz])
    s = Solid([_cone_leaf(rmax1, rmax2, hz)], [(0, 0)], -h, h, name)
    if rmin1 > 0.0 or rmin2 > 0.0:
        inner = Solid([_cone_leaf(rmin1, rmin2, hz)], [(0, 0)], -h, h)
        s = boolean("sub", s, inner, name=name)
    if dphi < 2 * np.pi - 1e-12:
        s = boolean("inter", s, _phi_wedge(phi0, dphi, rmax, hz), name=name)
    s.lo, s.hi = -h, h
    s.params = dict(rmin1=rmin1, rmax1=rmax1, rmin2=rmin2, rmax2=rmax2, hz=hz, phi0=phi0, dphi=dphi)
    return s


def polycone(z, rmin, rmax, name="polycone") -> Solid:
    """Return G4Polycone with a full turn: the union of the frusta between consecutive z planes (radii [rmin_k,
    rmax_k] at z_k, linear between), each a cone leaf (less its inner cone) on its own slab."""  # noqa: D205, D209
    z, rmin, rmax = (np.asarray(v, np.float64) for v in (z, rmin, rmax))
    if len(z) < 2 or np.any(np.diff(z) < 0) or len(rmin) != len(z) or len(rmax) != len(z):
        raise ValueError(f"polycone {name!r}: two or more z planes in order, one rmin and rmax each")
    parts = []
    for k in range(len(z) - 1):
        hz = 0.5 * (z[k + 1] - z[k])
        if hz <= 0.0:
            continue
        c = cons(rmin[k], rmax[k], rmin[k + 1], rmax[k + 1], hz, name=name)
        parts.append((c, 0.5 * (z[k] + z[k + 1])))
    s = None
    for c, zc in parts:
        s = _moved(c, zc, name) if s is None else boolean("union", s, c, np.eye(3), np.array([0.0, 0.0, zc]),
                                                           name=name)
    r = float(np.max(rmax))
    s.lo, s.hi = np.array([-r, -r, z[0]]), np.array([r, r, z[-1]])
    s.params = dict(z=z.tolist(), rmin=rmin.tolist(), rmax=rmax.tolist())
    return s


def _moved(s: Solid, zc: float, name: str) -> Solid:
    """Return the solid translated by zc along z."""
    t = np.array([0.0, 0.0, zc])
    return Solid([lf.moved(np.eye(3), t) for lf in s.leaves], list(s.program), s.lo + t, s.hi + t, name)


def ellipsoid(ax: float, by: float, cz: float, zcut1: float = -np.inf, zcut2: float = np.inf,
              name="ellipsoid") -> Solid:
    """Return G4Ellipsoid: (x/ax)^2 + (y/by)^2 + (z/cz)^2 <= 1 cut to zcut1 <= z <= zcut2 (one leaf)."""
    if not (ax > 0 and by > 0 and cz > 0):
        raise ValueError(f"ellipsoid {name!r}: positive semi-axes")
    z1, z2 = max(-cz, zcut1), min(cz, zcut2)
    N, H = [], []
    if z2 < cz:
        N.append([0, 0, 1.0]), H.append(z2)
    if z1 > -cz:
        N.append([0, 0, -1.0]), H.append(-z1)
    lf = Leaf(1, _z3(), _z3(), np.diag([1 / ax**2, 1 / by**2, 1 / cz**2]), _z3(), -1.0,
              np.asarray(N, np.float64).reshape(-1, 3), np.asarray(H, np.float64))
    return Solid([lf], [(0, 0)], np.array([-ax, -by, z1]), np.array([ax, by, z2]), name,
                 dict(ax=ax, by=by, cz=cz, zcut1=z1, zcut2=z2))


def trd(x1: float, x2: float, y1: float, y2: float, hz: float, name="trd") -> Solid:
    """Return G4Trd: half-lengths (x1, y1) at z = -hz and (x2, y2) at z = +hz, linear between (six planes)."""
    if not (x1 >= 0 and x2 >= 0 and y1 >= 0 and y2 >= 0 and hz > 0 and x1 + x2 > 0 and y1 + y2 > 0):
        raise ValueError(f"trd {name!r}: non-negative half-lengths, hz > 0")
    # the side x <= xm + sx z (xm the mid half-length, sx its slope): (1, 0, -sx).(x, y, z) <= xm
    xm, sx = 0.5 * (x1 + x2), (x2 - x1) / (2 * hz)
    ym, sy = 0.5 * (y1 + y2), (y2 - y1) / (2 * hz)
    N = np.array([[1, 0, -sx], [-1, 0, -sx], [0, 1, -sy], [0, -1, -sy], [0, 0, 1.0], [0, 0, -1.0]], np.float64)
    H = np.array([xm, xm, ym, ym, hz, hz], np.float64)
    nn = np.linalg.norm(N, axis=1)
    lf = Leaf(0, _z3(), _z3(), np.zeros((3, 3)), _z3(), 0.0, N / nn[:, None], H / nn)
    m = np.array([max(x1, x2), max(y1, y2), hz])
    return Solid([lf], [(0, 0)], -m, m, name, dict(x1=x1, x2=x2, y1=y1, y2=y2, hz=hz))


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
