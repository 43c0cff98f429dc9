"""The reference's navigator edge-case scenes (tests/unit/nav_toys.py:51-150: nest, bool, wires, coincident and its
"near" variant a few bands apart), restated as detectors of this build: definitions place daughters (R, t) in
their mother's frame (x_mother = R x + t), an assembly (no solid) places its daughters in its own mother, lengths
are GDML's (a box's full edges, a tube's full length; metres).

Every medium has the one index 1.33 (no refraction: the paths are the geometry's); a class per medium; every placed
volume other than the world is a sensor whose entry is a transmitting mirror R = 0, T = 0.9, eff = 1 (a tenth
detected, the rest straight on), every other crossing between media the matched interface, the world's exit an
absorber.  Not restated: the split union (its 21 leaves are more than a solid holds here), overlapping siblings and
the tetrahedron mesh (outside this navigator)."""

from __future__ import annotations

import numpy as np

from lumen_amd.detector import navarrays, solids
from lumen_amd.detector.model import Detector, Medium, SensorGroup, SurfaceRow
from lumen_amd.params import Params, Table

E_KNOTS = np.array([1.5, 5.0])


def rz(deg):
    a = np.radians(deg)
    return np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1.0]])


def rx(deg):
    a = np.radians(deg)
    return np.array([[1.0, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])


def box(x, y, z):
    return solids.box(0.5 * x, 0.5 * y, 0.5 * z)


def nest():
    """boxes and spheres nested 3 deep, off centre, rotated; a daughter flush with its mother's face; siblings sharing
    a face"""
    return {
        "world": (box(2.0, 2.0, 2.0), 1, [("tank", rz(20), (0.1, 0, 0)), ("left", np.eye(3), (-0.7, 0, 0)),
                                          ("right", np.eye(3), (-0.5, 0, 0))]),
        "tank": (box(0.8, 0.9, 1.0), 2, [("ball", np.eye(3), (-0.05, 0.05, 0)), ("flush", np.eye(3), (0.3, 0, 0))]),
        "ball": (solids.orb(0.2), 3, [("core", np.eye(3), (0.03, 0, 0))]),
        "core": (solids.orb(0.1), 4, []),
        "flush": (box(0.2, 0.1, 0.1), 5, []),
        "left": (box(0.2, 0.2, 0.2), 6, []),
        "right": (box(0.2, 0.2, 0.2), 7, []),
    }


def boolean():
    """a box minus a sphere, a union of two boxes (one solid), and an assembly placing two cubes"""
    sub = solids.boolean("sub", box(0.4, 0.4, 0.4), solids.orb(0.15), t=(0.1, 0, 0))
    uni = solids.boolean("union", box(0.3, 0.1, 0.1), box(0.1, 0.3, 0.1), R=rz(30), t=(0.05, 0, 0))
    return {
        "world": (box(2.0, 2.0, 2.0), 1, [("sub", np.eye(3), (-0.4, 0, 0)), ("uni", rx(10), (0.4, 0, 0)),
                                          ("asm", rz(15), (0, 0.5, 0))]),
        "sub": (sub, 2, []), "uni": (uni, 3, []),
        "asm": (None, 0, [("cube", np.eye(3), (-0.1, 0, 0)), ("cube", np.eye(3), (0.1, 0, 0))]),
        "cube": (box(0.1, 0.1, 0.1), 4, []),
    }


def wires(n=40):
    """a plane volume of n thin wires (r 0.5 mm, 600 mm long) all at 60 degrees to the plane's z axis: a frame
    group"""
    pls = [("wire", rx(-60), (0.0, 0.0, -0.3 + 0.6 * (k + 0.5) / n)) for k in range(n)]
    return {
        "world": (box(2.0, 2.0, 2.0), 1, [("plane", np.eye(3), (0, 0, 0))]),
        "plane": (box(0.02, 1.6, 1.6), 1, pls),
        "wire": (solids.tube(0.0005, 0.3), 2, []),
    }


def coincident(gap=0.0):
    """coincident faces (``gap`` 0), or faces a few bands apart, between media shared and not: env (the world's
    medium) holds glass flush with three of its faces, glass holds inner flush with the same faces; edge and edgein
    (one medium) flush with the world's +x face; s1 and s2 (one medium) sharing a face, s2's daughter flush with it"""
    g = gap
    return {
        "world": (box(2.0, 2.0, 2.0), 1, [("env", np.eye(3), (0, 0, 0)), ("edge", np.eye(3), (0.8 - g, 0, 0)),
                                          ("s1", np.eye(3), (0, 0.8, 0)), ("s2", np.eye(3), (0.2, 0.8, 0))]),
        "env": (box(0.6, 0.6, 0.6), 1, [("glass", np.eye(3), (0, 0, 0.15 - g))]),
        "glass": (box(0.6 - 2 * g, 0.6 - 2 * g, 0.3), 2, [("inner", np.eye(3), (0, 0, 0.1 - g))]),
        "inner": (box(0.6 - 4 * g, 0.6 - 4 * g, 0.1), 3, []),
        "edge": (box(0.4, 0.4, 0.4), 4, [("edgein", np.eye(3), (0.1 - g, 0, 0))]),
        "edgein": (box(0.2, 0.4, 0.4), 4, []),
        "s1": (box(0.2, 0.2, 0.2), 5, []),
        "s2": (box(0.2, 0.2, 0.2), 5, [("s2in", np.eye(3), (-0.05 + g, 0, 0))]),
        "s2in": (box(0.1, 0.1, 0.1), 6, []),
    }


SCENES = {"nest": nest, "bool": boolean, "wires": wires, "coincident": coincident,
          "near": lambda: coincident(3e-7)}


def detector(name: str) -> Detector:
    """Return the scene ``name`` as a detector (module docstring)."""
    defs = SCENES[name]()
    vols, kinds, kind_of = [], [], {}

    def place(d, R, t, parent, label):
        solid, med, pls = defs[d]
        me = parent
        if solid is not None:
            if d not in kind_of:
                kind_of[d] = len(kinds)
                kinds.append(solid)
            me = len(vols)
            vols.append(navarrays.Volume(label, kind_of[d], R, t, parent, med))
        for k, (child, Rc, tc) in enumerate(pls):
            place(child, R @ Rc, R @ np.asarray(tc, float) + t, me, f"{label}/{child}_{k}")

    place("world", np.eye(3), np.zeros(3), -1, "world")
    n_med = 1 + max(v.cls for v in vols)
    out = n_med
    nav = navarrays.build(vols, kinds, outside_cls=out)
    C = n_med + 1
    media, values = [], {}
    for m in range(n_med):
        key = f"materials/M{m}/RINDEX"
        values[key] = Table(E_KNOTS, np.array([1.33, 1.33]))
        media.append(Medium(f"M{m}", key))
    rows = [SurfaceRow("absorber", "ABSORBER"), SurfaceRow("interface", "DIELECTRIC"),
            SurfaceRow("sensor", "SPECULAR_METAL", {"R": 0.0, "T": 0.9, "eff": 1.0}, transmitting=True)]
    row_of = np.full((C, C), 1, np.int32)
    row_of[:, out] = 0
    np.fill_diagonal(row_of, -1)
    sensor_cls = {v.cls for v in vols[1:]}
    for c in sensor_cls:  # entries into a class of sensor volumes
        row_of[[a for a in range(n_med) if a != c], c] = 2
    sensor_of_vol = np.full(len(vols) + 1, -1, np.int32)
    sensor_of_vol[1:len(vols)] = np.arange(len(vols) - 1)
    names = [v.name for v in vols[1:]]
    return Detector(f"scene_{name}", nav, media, np.arange(C, dtype=np.int32) - (np.arange(C) == out) * (out + 1),
                    [f"M{m}" for m in range(n_med)] + ["OUT"], rows, row_of, sensor_of_vol, names,
                    [SensorGroup("vols", np.arange(len(names)))], Params(values))
