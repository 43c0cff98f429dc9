"""A GDML-subset reader and the detector builder behind ``lumen_amd.load`` (written for this engine).

It reads what the tutorial tank, the 1,000-PMT water tank and the LAr box use: ``define`` (``matrix``, ``constant``, ``position``, ``rotation``),
``materials`` (material properties by matrix reference: RINDEX, ABSLENGTH, RAYLEIGH, the Mie MIEHG constants, the
wavelength shifter's WLS properties), ``solids`` (``box``, ``tube`` (a phi cut too), ``cone``, ``polycone`` without a
phi cut, ``orb``, ``sphere`` with polar cuts, ``ellipsoid``, ``trd``, ``torus``, the booleans, ``opticalsurface``),
``structure`` (``volume``, ``physvol``,
``bordersurface``, ``skinsurface``) and ``setup``.  A recipe names the detector volume (light leaving it is absorbed)
and the sensor groups (counted across a border into each placement of a logical volume), as the reference's
``lumen.Recipe`` does (``examples/example_tank.py``).

Geant4 conventions kept: a GDML physvol rotation is the inverse of the placement's frame rotation (G4GDMLRead); a
table's energies are in eV and lengths in m; an optical surface maps to the engine's surface types as
``lumen_amd.detector.model`` lists them (G4OpBoundaryProcess's finishes and types).
"""

from __future__ import annotations

import json
import math
import re
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

from ..params import Params, Table
from . import navarrays, solids
from .model import G4_COLUMN, Detector, Medium, SensorGroup, SurfaceRow

UNITS = {"mm": 1e-3, "cm": 1e-2, "m": 1.0, "km": 1e3, "um": 1e-6, "nm": 1e-9, "eV": 1.0, "keV": 1e3, "MeV": 1e6,
         "rad": 1.0, "deg": math.pi / 180.0, "ns": 1.0, "s": 1e9}


def _num(s: str, consts: dict) -> float:
    """Return a GDML number: a float, a constant, or ``x*unit``."""
    s = s.strip()
    if "*" in s:
        a, u = s.split("*", 1)
        return _num(a, consts) * UNITS[u.strip()]
    if s in consts:
        return consts[s]
    return float(s)


def _rot(x, y, z):
    """Return the frame rotation of a GDML physvol rotation (x, y, z) in radians: the inverse of Rx Ry Rz."""
    cx, sx, cy, sy, cz, sz = math.cos(x), math.sin(x), math.cos(y), math.sin(y), math.cos(z), math.sin(z)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return (Rx @ Ry @ Rz).T


@dataclass
class Recipe:
    """Which volume is the detector and which placements are sensors (the reference's ``lumen.Recipe``)."""

    name: str
    detector_volume: str
    groups: tuple = ()


@dataclass
class Group:
    """A sensor group of a recipe: the placements of a logical volume ``lv``; ``values``: a sensor value set.

    ``rule`` (the reference's ``rules/sensors.py``): "border_into", counted at the optical surfaces bordering into
    the volume (their EFFICIENCY); "any_entry", a sensitive volume: every photon entering it from any side is absorbed
    and counted with ``entry_efficiency`` (1 with a value set, whose efficiency acts after tracking; else the setting,
    default 1), and with ``count`` "volume" (Geant4's sensitive-detector convention, the default) a photon emitted
    inside it is counted at its emission with the entry's detection ("entry": it leaves uncounted).  Leaving a sensor
    volume through a border surface that detects (an exit sensor) credits the volume left, whatever the rule.
    """

    name: str
    lv: str
    values: str | None = None
    rule: str = "border_into"
    count: str = "volume"
    entry_efficiency: float | None = None

    def __post_init__(self):
        if self.rule not in ("border_into", "any_entry"):
            raise ValueError(f"sensor group {self.name!r}: rule {self.rule!r} is not border_into or any_entry")
        if self.count not in ("volume", "entry"):
            raise ValueError(f"sensor group {self.name!r}: count {self.count!r} is not volume or entry")


@dataclass
class _Doc:
    consts: dict = field(default_factory=dict)
    matrices: dict = field(default_factory=dict)
    positions: dict = field(default_factory=dict)
    rotations: dict = field(default_factory=dict)
    materials: dict = field(default_factory=dict)
    solids: dict = field(default_factory=dict)
    osurfaces: dict = field(default_factory=dict)
    volumes: dict = field(default_factory=dict)
    borders: list = field(default_factory=list)
    skins: list = field(default_factory=list)
    world: str = ""


def _lunit(el, default="mm"):
    return UNITS[el.get("lunit", default)]


def _aunit(el, default="rad"):
    return UNITS[el.get("aunit", default)]


def _parse(path: str) -> _Doc:
    root = ET.parse(path).getroot()
    d = _Doc()
    for el in root.find("define") or []:
        if el.tag == "constant":
            d.consts[el.get("name")] = _num(el.get("value"), d.consts)
        elif el.tag == "matrix":
            vals = [_num(v, d.consts) for v in el.get("values").split()]
            d.matrices[el.get("name")] = np.asarray(vals).reshape(-1, int(el.get("coldim", "1")))
        elif el.tag == "position":
            u = UNITS[el.get("unit", "mm")]
            d.positions[el.get("name")] = np.array([_num(el.get(a, "0"), d.consts) * u for a in "xyz"])
        elif el.tag == "rotation":
            u = UNITS[el.get("unit", "rad")]
            d.rotations[el.get("name")] = _rot(*[_num(el.get(a, "0"), d.consts) * u for a in "xyz"])
    for el in root.find("materials") or []:
        if el.tag == "material":
            d.materials[el.get("name")] = {p.get("name"): p.get("ref") for p in el.findall("property")}
    for el in root.find("solids") or []:
        name = el.get("name")
        if el.tag == "box":
            u = _lunit(el)
            d.solids[name] = solids.box(*[0.5 * _num(el.get(a), d.consts) * u for a in "xyz"], name=name)
        elif el.tag == "tube":
            u, au = _lunit(el), _aunit(el)
            dphi = _num(el.get("deltaphi"), d.consts) * au if el.get("deltaphi") else 2 * math.pi
            rmax, hz = _num(el.get("rmax"), d.consts) * u, 0.5 * _num(el.get("z"), d.consts) * u
            rmin = _num(el.get("rmin", "0"), d.consts) * u
            if abs(dphi - 2 * math.pi) > 1e-9:  # a phi cut: the cone of equal radii with its wedge
                d.solids[name] = solids.cons(rmin, rmax, rmin, rmax, hz, _num(el.get("startphi", "0"), d.consts) * au,
                                             dphi, name=name)
            else:
                d.solids[name] = solids.tube(rmax, hz, rmin, name=name)
        elif el.tag == "cone":
            u, au = _lunit(el), _aunit(el)
            dphi = _num(el.get("deltaphi"), d.consts) * au if el.get("deltaphi") else 2 * math.pi
            d.solids[name] = solids.cons(*[_num(el.get(k, "0"), d.consts) * u for k in ("rmin1", "rmax1", "rmin2",
                                                                                      "rmax2")],
                                         0.5 * _num(el.get("z"), d.consts) * u,
                                         _num(el.get("startphi", "0"), d.consts) * au, dphi, name=name)
        elif el.tag == "torus":
            u, au = _lunit(el), _aunit(el)
            dphi = _num(el.get("deltaphi"), d.consts) * au if el.get("deltaphi") else 2 * math.pi
            d.solids[name] = solids.torus(*[_num(el.get(k, "0"), d.consts) * u for k in ("rmin", "rmax", "rtor")],
                                          _num(el.get("startphi", "0"), d.consts) * au, dphi, name=name)
        elif el.tag == "polycone":
            u, au = _lunit(el), _aunit(el)
            dphi = _num(el.get("deltaphi"), d.consts) * au if el.get("deltaphi") else 2 * math.pi
            if abs(dphi - 2 * math.pi) > 1e-9:
                raise ValueError(f"gdml: polycone {name!r} with a phi cut is not supported")
            zp = el.findall("zplane")
            d.solids[name] = solids.polycone([_num(z.get("z"), d.consts) * u for z in zp],
                                             [_num(z.get("rmin", "0"), d.consts) * u for z in zp],
                                             [_num(z.get("rmax"), d.consts) * u for z in zp], name=name)
        elif el.tag == "ellipsoid":
            u = _lunit(el)
            d.solids[name] = solids.ellipsoid(
                *[_num(el.get(k), d.consts) * u for k in ("ax", "by", "cz")],
                _num(el.get("zcut1"), d.consts) * u if el.get("zcut1") else -math.inf,
                _num(el.get("zcut2"), d.consts) * u if el.get("zcut2") else math.inf, name=name)
        elif el.tag == "trd":
            u = _lunit(el)
            d.solids[name] = solids.trd(*[0.5 * _num(el.get(k), d.consts) * u for k in ("x1", "x2", "y1", "y2", "z")],
                                        name=name)
        elif el.tag == "orb":
            d.solids[name] = solids.orb(_num(el.get("r"), d.consts) * _lunit(el), name=name)
        elif el.tag == "sphere":
            u, au = _lunit(el), _aunit(el)
            dphi = _num(el.get("deltaphi", "0"), d.consts) * au if el.get("deltaphi") else 2 * math.pi
            if abs(dphi - 2 * math.pi) > 1e-9:
                raise ValueError(f"gdml: sphere {name!r} with a phi cut is not supported")
            t0 = _num(el.get("starttheta", "0"), d.consts) * au
            dt = _num(el.get("deltatheta"), d.consts) * au if el.get("deltatheta") else math.pi
            d.solids[name] = solids.sphere(_num(el.get("rmin", "0"), d.consts) * u, _num(el.get("rmax"), d.consts) * u,
                                           t0, dt, name=name)
        elif el.tag in ("union", "subtraction", "intersection"):
            a, b = d.solids[el.find("first").get("ref")], d.solids[el.find("second").get("ref")]
            pos = el.find("position")
            posr = el.find("positionref")
            rot = el.find("rotation")
            t = (d.positions[posr.get("ref")] if posr is not None else
                 np.array([_num(pos.get(k, "0"), d.consts) * UNITS[pos.get("unit", "mm")] for k in "xyz"])
                 if pos is not None else np.zeros(3))
            R = (_rot(*[_num(rot.get(k, "0"), d.consts) * UNITS[rot.get("unit", "rad")] for k in "xyz"])
                 if rot is not None else np.eye(3))
            op = {"union": "union", "subtraction": "sub", "intersection": "inter"}[el.tag]
            d.solids[name] = solids.boolean(op, a, b, R, t, name=name)
        elif el.tag == "opticalsurface":
            d.osurfaces[name] = dict(model=el.get("model", "glisur"), finish=el.get("finish", "polished"),
                                     type=el.get("type", "dielectric_dielectric"),
                                     value=_num(el.get("value", "1.0"), d.consts),
                                     props={p.get("name"): p.get("ref") for p in el.findall("property")})
        else:
            raise ValueError(f"gdml: solid <{el.tag}> is not supported")
    for el in root.find("structure") or []:
        if el.tag == "volume":
            pvs = []
            for pv in el.findall("physvol"):
                vref = pv.find("volumeref").get("ref")
                pos, posr = pv.find("position"), pv.find("positionref")
                rot, rotr = pv.find("rotation"), pv.find("rotationref")
                t = (d.positions[posr.get("ref")] if posr is not None else
                     np.array([_num(pos.get(k, "0"), d.consts) * UNITS[pos.get("unit", "mm")] for k in "xyz"])
                     if pos is not None else np.zeros(3))
                R = (d.rotations[rotr.get("ref")] if rotr is not None else
                     _rot(*[_num(rot.get(k, "0"), d.consts) * UNITS[rot.get("unit", "rad")] for k in "xyz"])
                     if rot is not None else np.eye(3))
                pvs.append(dict(name=pv.get("name", vref), lv=vref, R=R, t=t))
            d.volumes[el.get("name")] = dict(material=el.find("materialref").get("ref"),
                                             solid=el.find("solidref").get("ref"), pvs=pvs)
        elif el.tag == "bordersurface":
            refs = [p.get("ref") for p in el.findall("physvolref")]
            d.borders.append((el.get("surfaceproperty"), refs[0], refs[1]))
        elif el.tag == "skinsurface":
            d.skins.append((el.get("surfaceproperty"), el.find("volumeref").get("ref")))
    d.world = root.find("setup").find("world").get("ref")
    return d


def _surface_row(name, s: dict):
    """Return (the engine's surface row of a GDML optical surface, the defaults it adds as parameters) by
    G4OpBoundaryProcess's types and finishes (the reference's ``rules/optics.py:28-50`` table): a dielectric_metal
    reads REFLECTIVITY at Geant4's default 1 when the surface omits it (a parameter at that value), a complex index
    (REALRINDEX, IMAGINARYRINDEX) makes a photocathode, TRANSMITTANCE a transmitting row."""  # noqa: D205, D209
    props = s["props"]
    cols = {G4_COLUMN[k]: f"surfaces/{name}/{k}" for k in props if k in G4_COLUMN}
    defaults = {}
    typ, fin, model = s["type"], s["finish"], s["model"]
    trans = "T" in cols
    if typ == "dielectric_metal":
        if "R" not in cols and not ("nr" in cols and "nk" in cols):
            cols["R"] = f"surfaces/{name}/REFLECTIVITY"
            defaults[cols["R"]] = np.float64(1.0)
        if "nr" in cols and "nk" in cols:
            if not (fin == "polished" or (model == "glisur" and float(s["value"]) >= 1.0)):
                raise ValueError(f"gdml: optical surface {name!r}: a complex index on a ground finish (the facet's "
                                 "Fresnel) is not supported")
            cols.pop("R", None)
            return SurfaceRow(name, "PHOTOCATHODE", cols, transmitting=trans), defaults
        if model == "glisur" and fin == "ground" and float(s["value"]) < 1.0:  # the facet path at 1 - polish
            cols["polish"] = f"surfaces/{name}/polish"
            defaults[cols["polish"]] = np.float64(s["value"])
            return SurfaceRow(name, "GLISUR_METAL", cols, transmitting=trans), defaults
        if fin == "polished" or model == "glisur":  # glisur: any other finish, or polish 1, is the mirror
            return SurfaceRow(name, "SPECULAR_METAL", cols, transmitting=trans), defaults
        if fin == "ground" and model == "unified":
            if float(s["value"]) == 0.0:  # sigma_alpha 0: GetFacetNormal returns the normal
                return SurfaceRow(name, "UNIFIED_METAL_NORMAL", cols, transmitting=trans), defaults
            cols["sigma_alpha"] = f"surfaces/{name}/sigma_alpha"
            return SurfaceRow(name, "UNIFIED_METAL", cols, transmitting=trans), defaults
    if typ == "dielectric_dielectric":
        if fin in ("groundfrontpainted", "polishedfrontpainted"):
            if "R" not in cols:
                cols["R"] = f"surfaces/{name}/REFLECTIVITY"
                defaults[cols["R"]] = np.float64(1.0)
            cols["p_spike"] = 1.0 if fin == "polishedfrontpainted" else 0.0
            return SurfaceRow(name, "PAINTED_T" if trans else "PAINTED", cols), defaults
        if fin == "polished" or (model == "glisur" and float(s["value"]) >= 1.0):
            return SurfaceRow(name, "DIELECTRIC_RT" if "R" in cols or "T" in cols else "DIELECTRIC", cols), defaults
        if model == "glisur" and fin == "ground":  # the facet path reflecting or refracting at each facet
            if "R" not in cols:
                cols["R"] = f"surfaces/{name}/REFLECTIVITY"
                defaults[cols["R"]] = np.float64(1.0)
            cols["polish"] = f"surfaces/{name}/polish"
            defaults[cols["polish"]] = np.float64(s["value"])
            return SurfaceRow(name, "GLISUR_DIELECTRIC", cols), defaults
    raise ValueError(f"gdml: optical surface {name!r} ({model}, {typ}, {fin}) is not supported")


def _table(d: _Doc, ref: str) -> Table:
    m = d.matrices[ref]
    if m.shape[1] != 2:
        raise ValueError(f"gdml: property matrix {ref!r} is not (energy, value) pairs")
    return Table(m[:, 0].copy(), m[:, 1].copy())


def load(path: str, recipe: Recipe) -> Detector:
    """Return the Detector of a GDML file under ``recipe`` (see the module docstring)."""
    d = _parse(path)
    # the detector volume's first placement: walk the world's tree, depth first, in physvol order
    placed = []  # (pv name, lv, R world, t world, parent index)

    def walk(lv, R, t, parent, pvname, inside):
        me = None
        if inside or lv == recipe.detector_volume:
            me = len(placed)
            placed.append((pvname, lv, R, t, parent))
            inside = True
        for pv in d.volumes[lv]["pvs"]:
            walk(pv["lv"], R @ pv["R"], R @ pv["t"] + t, me if inside else -1, pv["name"], inside)

    walk(d.world, np.eye(3), np.zeros(3), -1, d.world, False)
    if not placed:
        raise ValueError(f"gdml: detector volume {recipe.detector_volume!r} is not placed")
    kinds, kind_of = [], {}
    for _, lv, _, _, _ in placed:
        s = d.volumes[lv]["solid"]
        if s not in kind_of:
            kind_of[s] = len(kinds)
            kinds.append(d.solids[s])
    # media: one per material met (in order of first placement); the outside is opaque
    mats = []
    for _, lv, _, _, _ in placed:
        m = d.volumes[lv]["material"]
        if m not in mats:
            mats.append(m)
    values, media = {}, []
    for m in mats:
        props = d.materials.get(m, {})
        keys = {}
        for p in ("RINDEX", "ABSLENGTH", "RAYLEIGH", "MIEHG", "WLSABSLENGTH", "WLSCOMPONENT"):
            if p in props:
                values[f"materials/{m}/{p}"] = _table(d, props[p])
                keys[p] = f"materials/{m}/{p}"
        for p in ("MIEHG_FORWARD", "MIEHG_BACKWARD", "MIEHG_FORWARD_RATIO", "WLSMEANNUMBERPHOTONS", "WLSTIMECONSTANT"):
            if p in props:  # constant properties: one value
                values[f"materials/{m}/{p}"] = np.float64(d.matrices[props[p]].reshape(-1)[-1])
                keys[p] = f"materials/{m}/{p}"
        ray = keys.get("RAYLEIGH")
        if ray is None and m.split("0x")[0] == "Water" and "RINDEX" in keys:
            ray = "water_rule"
        extra = {}
        if "MIEHG" in keys:
            extra.update(mie_length=keys["MIEHG"], mie_forward=keys.get("MIEHG_FORWARD", 0.0),
                         mie_backward=keys.get("MIEHG_BACKWARD", 0.0), mie_ratio=keys.get("MIEHG_FORWARD_RATIO", 1.0))
        if "WLSABSLENGTH" in keys:
            extra.update(wls_length=keys["WLSABSLENGTH"], wls_spectrum=keys.get("WLSCOMPONENT"),
                         wls_yield=keys.get("WLSMEANNUMBERPHOTONS", 1.0), wls_tau=keys.get("WLSTIMECONSTANT", 0.0))
        media.append(Medium(m, keys.get("RINDEX"), keys.get("ABSLENGTH"), ray, **extra))
    # surfaces: border (pv1 -> pv2) and skin (either way); classes by (material, surfaces entering it)
    pv_index = {p[0]: i for i, p in enumerate(placed)}
    srows, srow_of = [], {}
    for sname, s in d.osurfaces.items():
        srow_of[sname] = len(srows)
        row, defaults = _surface_row(sname, s)
        srows.append(row)
        values.update(defaults)
        for prop, ref in s["props"].items():
            if prop in G4_COLUMN:
                values[f"surfaces/{sname}/{prop}"] = _table(d, ref)
        if s["type"] == "dielectric_metal" and s["finish"] == "ground" and s["model"] == "unified":
            values[f"surfaces/{sname}/sigma_alpha"] = np.float64(s["value"])
    border_in, border_out = {}, {}
    for sname, a, b in d.borders:
        if a in pv_index and b in pv_index:
            border_in.setdefault(pv_index[b], []).append((sname, pv_index[a]))
            border_out.setdefault(pv_index[a], []).append(sname)
    skin_lv = {lv: sname for sname, lv in d.skins}
    group_of_lv = {g.lv: g for g in recipe.groups}
    cls_key, cls_names, vol_cls = {}, [], []
    for i, (pvname, lv, _, _, _) in enumerate(placed):  # a class: medium, borders in and out, skin, sensor group
        g = group_of_lv.get(lv)
        key = (d.volumes[lv]["material"], tuple(sorted(s for s, _ in border_in.get(i, []))), skin_lv.get(lv),
               None if g is None else g.name, tuple(sorted(border_out.get(i, []))))
        if key not in cls_key:
            cls_key[key] = len(cls_names)
            cls_names.append(f"{key[0]}" + (f"|{','.join(key[1])}" if key[1] else "")
                             + (f"|skin:{key[2]}" if key[2] else "") + (f"|sensor:{key[3]}" if key[3] else "")
                             + (f"|out:{','.join(key[4])}" if key[4] else ""))
        vol_cls.append(cls_key[key])
    out_cls = len(cls_names)
    cls_names.append("OUT")
    C = len(cls_names)
    class_medium = np.full(C, -1, np.int32)
    for key, c in cls_key.items():
        m = mats.index(key[0])
        class_medium[c] = m if media[m].rindex is not None else -1
    absorber = len(srows)
    srows.append(SurfaceRow("absorber", "ABSORBER"))
    dielectric = None
    row_of = np.full((C, C), -1, np.int32)
    for a in range(C):
        for b in range(C):
            if a == b:
                continue
            if class_medium[a] < 0:  # no photon travels in an opaque medium
                continue
            if b == out_cls or class_medium[b] < 0:
                row_of[a, b] = absorber
            else:
                if dielectric is None:
                    dielectric = len(srows)
                    srows.append(SurfaceRow("dielectric", "DIELECTRIC"))
                row_of[a, b] = dielectric
    def put(a, b, r):  # an interface into a medium without RINDEX: Geant4 kills the photon before the chain
        interface = srows[r].type in ("DIELECTRIC", "DIELECTRIC_RT", "GLISUR_DIELECTRIC")
        row_of[a, b] = absorber if interface and (b == out_cls or class_medium[b] < 0) else r

    for i, lst in border_in.items():
        for sname, j in lst:
            put(vol_cls[j], vol_cls[i], srow_of[sname])
    for i, (_, lv, _, _, _) in enumerate(placed):
        if lv in skin_lv:
            r = srow_of[skin_lv[lv]]
            p = placed[i][4]
            if p >= 0:
                put(vol_cls[p], vol_cls[i], r)
                put(vol_cls[i], vol_cls[p], r)
    born_row = np.full(C, -1, np.int32)
    for g in recipe.groups:  # a sensitive volume: every entry counted (a counting row), a photon born inside too
        if g.rule != "any_entry":
            continue
        eff = 1.0 if g.values is not None else (1.0 if g.entry_efficiency is None else float(g.entry_efficiency))
        r = len(srows)
        srows.append(SurfaceRow(f"{g.name}/entry", "COUNTING", {"eff": eff}))
        for sc in {vol_cls[i] for i, p in enumerate(placed) if p[1] == g.lv}:
            for a in range(C):
                if a != sc and a != out_cls and class_medium[a] >= 0:
                    row_of[a, sc] = r
            if g.count == "volume":
                born_row[sc] = r
    vols = [navarrays.Volume(p[0], kind_of[d.volumes[p[1]]["solid"]], p[2], p[3], p[4], vol_cls[i])
            for i, p in enumerate(placed)]
    nav, _ = navarrays.cached(vols, kinds, out_cls)
    sensor_of_vol = np.full(len(vols) + 1, -1, np.int32)
    names, groups = [], []
    for g in recipe.groups:
        ids = []
        for i, p in enumerate(placed):
            if p[1] == g.lv:
                sensor_of_vol[i] = len(names)
                ids.append(len(names))
                names.append("/".join(_path(placed, i)))
        qe = eff = ce = None
        if g.values is not None:
            with open(g.values) as f:
                vs = json.load(f)
            if vs.get("qe"):
                values[f"sensors/{g.name}/qe"] = Table(np.asarray(vs["qe"]["wavelength_nm"], float),
                                                       np.asarray(vs["qe"]["values"], float))
                qe = f"sensors/{g.name}/qe"
            if vs.get("ce"):  # CE over the cathode position angle: theta_deg knots in the file, radians here
                values[f"sensors/{g.name}/ce"] = Table(np.radians(np.asarray(vs["ce"]["theta_deg"], float)),
                                                       np.asarray(vs["ce"]["values"], float))
                ce = f"sensors/{g.name}/ce"
            values[f"sensors/{g.name}/efficiency"] = np.float64(vs.get("efficiency", 1.0))
            eff = f"sensors/{g.name}/efficiency"
        groups.append(SensorGroup(g.name, np.asarray(ids, np.int64), qe, eff, ce))
    credit_left = np.zeros((C, C), bool)
    for sname, a, b in d.borders:  # an exit sensor: a detecting border out of a sensor volume credits the volume left
        if a in pv_index and b in pv_index and sensor_of_vol[pv_index[a]] >= 0:
            row = srows[srow_of[sname]]
            if "eff" in row.params() and row._fixed("eff") != 0.0:
                credit_left[vol_cls[pv_index[a]], vol_cls[pv_index[b]]] = True
    origin = {k: "detector file" for k in values}
    for g in groups:
        for k in (g.qe, g.efficiency, g.ce):
            if k is not None:
                origin[k] = "sensor model defaults"
    return Detector(recipe.name, nav, media, class_medium, cls_names, srows, row_of, sensor_of_vol, names, groups,
                    Params(values, origin), credit_left=credit_left if credit_left.any() else None,
                    born_row=born_row if (born_row >= 0).any() else None)


def _path(placed, i):
    out = []
    while i >= 0:
        out.append(placed[i][0])
        i = placed[i][4]
    return out[::-1]
