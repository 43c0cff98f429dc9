"""Synthetic detectors built directly as navigator arrays and optics (no GDML): the 1,000-PMT water tank.

The water tank: a water cylinder (R = 2.0 m, H = 3.0 m) whose liner (the boundary into the outside) is Geant4's
unified ground dielectric_metal (R = 0.10, SPECULARLOBECONSTANT 1, sigma_alpha = 0.1 rad), with 1,000 sphere-cap
PMTs (sphere radius 0.10 m, cap height 0.04 m, one CSG leaf each: the sphere cut by its base plane) on a regular grid:
600 on the barrel (10 rows of 60) and 200 on each endcap (the square grid points of spacing 0.2256 m nearest the
axis), each cap's base 3 mm inside the wall and its dome facing the axis.  Every PMT's cathode is a border surface
(R = 0, EFFICIENCY 1) counted into the PMT; QE 0.25 is applied after tracking.

Water: RINDEX on 16 knots uniform in photon energy over 2.066-4.133 eV (600-300 nm), a Sellmeier form fixed by
n(600 nm) = 1.33 and n(300 nm) = 1.36; ABSLENGTH on the same knots, 5 * 16^u m with u ~ U(0, 1) (seeded); RAYLEIGH
60 (lambda / 400 nm)^4 m on the same knots.
"""

from __future__ import annotations

import numpy as np

from .. import units
from ..params import Params, Table
from . import navarrays, solids
from .model import Detector, Medium, SensorGroup, SurfaceRow

TANK_R, TANK_HZ = 2.0, 1.5
CAP_R, CAP_H = 0.10, 0.04
WALL_GAP = 0.003
N_RINDEX_KNOTS = 16
LAM_MIN, LAM_MAX = 300.0, 600.0


def sellmeier(lam_nm, n_lo=1.33, lam_lo=600.0, n_hi=1.36, lam_hi=300.0):
    """Return n(lambda) = sqrt(1 + B lambda^2 / (lambda^2 - C)) through (lam_lo, n_lo) and (lam_hi, n_hi)."""
    a, b = n_lo**2 - 1.0, n_hi**2 - 1.0
    l1, l2 = lam_lo**2, lam_hi**2
    # a (l1 - C) / l1 = b (l2 - C) / l2  ->  C
    C = (a - b) / (a / l1 - b / l2)
    B = a * (l1 - C) / l1
    lam2 = np.asarray(lam_nm, np.float64) ** 2
    return np.sqrt(1.0 + B * lam2 / (lam2 - C))


def water_tables(seed: int = 1):
    """Return (RINDEX, ABSLENGTH, RAYLEIGH) Tables on 16 knots uniform in energy over 300-600 nm."""
    e = np.linspace(units.HC / LAM_MAX, units.HC / LAM_MIN, N_RINDEX_KNOTS)
    lam = units.HC / e
    u = np.random.default_rng(seed).uniform(0.0, 1.0, N_RINDEX_KNOTS)
    return (Table(e, sellmeier(lam)), Table(e.copy(), 5.0 * 16.0**u), Table(e.copy(), 60.0 * (lam / 400.0) ** 4))


def _frame(n_in):
    """Return a rotation whose third column is n_in (a right-handed frame)."""
    n = np.asarray(n_in, float) / np.linalg.norm(n_in)
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(a, n)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(n, e1)
    return np.column_stack([e1, e2, n])


def pmt_grid(n_barrel_rows=10, n_phi=60, n_endcap=200, spacing=0.2256):
    """Return (positions of the cap base centres on the wall, inward normals) of the 1,000 PMTs."""
    pos, nrm = [], []
    zs = np.linspace(-TANK_HZ, TANK_HZ, n_barrel_rows + 2)[1:-1]
    for row, z in enumerate(zs):
        for k in range(n_phi):
            ph = 2.0 * np.pi * (k + 0.5 * (row % 2)) / n_phi  # alternate rows staggered by half a step
            c, s = np.cos(ph), np.sin(ph)
            pos.append([TANK_R * c, TANK_R * s, z])
            nrm.append([-c, -s, 0.0])
    g = np.arange(-12, 13) * spacing
    xy = np.array([[x, y] for x in g for y in g])
    order = np.lexsort((xy[:, 1], xy[:, 0], np.round(np.hypot(xy[:, 0], xy[:, 1]), 9)))
    ring = xy[order[:n_endcap]]
    for zc, nz in ((TANK_HZ, -1.0), (-TANK_HZ, 1.0)):
        for x, y in ring:
            pos.append([x, y, zc])
            nrm.append([0.0, 0.0, nz])
    return np.asarray(pos), np.asarray(nrm)


def water_tank(seed: int = 1, reflectivity=0.10, sigma_alpha=0.1, qe=0.25) -> Detector:
    """Return the 1,000-PMT water tank (module docstring)."""
    kinds = [solids.tube(TANK_R, TANK_HZ, name="tank"), solids.sphere_cap(CAP_R, CAP_H, name="pmt_cap")]
    vols = [navarrays.Volume("tank", 0, np.eye(3), np.zeros(3), -1, 0)]
    pos, nrm = pmt_grid()
    z0 = CAP_R - CAP_H
    for i, (p, n) in enumerate(zip(pos, nrm)):
        R = _frame(n)
        face = p + WALL_GAP * n
        vols.append(navarrays.Volume(f"tank/pmt_{i}", 1, R, face - z0 * n, 0, 1))
    nav = navarrays.build(vols, kinds, outside_cls=2)
    rindex, absl, ray = water_tables(seed)
    values = {
        "materials/Water/RINDEX": rindex,
        "materials/Water/ABSLENGTH": absl,
        "materials/Water/RAYLEIGH": ray,
        "surfaces/Liner/REFLECTIVITY": np.float64(reflectivity),
        "surfaces/Liner/SPECULARLOBECONSTANT": np.float64(1.0),
        "surfaces/Liner/sigma_alpha": np.float64(sigma_alpha),
        "surfaces/Cathode/REFLECTIVITY": np.float64(0.0),
        "surfaces/Cathode/EFFICIENCY": np.float64(1.0),
        "sensors/pmt/qe": np.full(len(pos), float(qe)),
        "sensors/pmt/efficiency": np.float64(1.0),
    }
    media = [Medium("Water", "materials/Water/RINDEX", "materials/Water/ABSLENGTH", "materials/Water/RAYLEIGH")]
    rows = [SurfaceRow("Liner", "UNIFIED_METAL", {"R": "surfaces/Liner/REFLECTIVITY",
                                                  "p_lobe": "surfaces/Liner/SPECULARLOBECONSTANT",
                                                  "sigma_alpha": "surfaces/Liner/sigma_alpha"}),
            SurfaceRow("Cathode", "SPECULAR_METAL", {"R": "surfaces/Cathode/REFLECTIVITY",
                                                     "eff": "surfaces/Cathode/EFFICIENCY"}),
            SurfaceRow("absorber", "ABSORBER")]
    row_of = np.full((3, 3), -1, np.int32)
    row_of[0, 2] = 0  # water -> outside: the liner
    row_of[0, 1] = 1  # water -> PMT: the cathode
    sensor_of_vol = np.full(nav.n + 1, -1, np.int32)
    sensor_of_vol[1: nav.n] = np.arange(nav.n - 1)
    names = [v.name for v in vols[1:]]
    groups = [SensorGroup("pmt", np.arange(len(names)), "sensors/pmt/qe", "sensors/pmt/efficiency")]
    return Detector("water_tank_1000", nav, media, np.array([0, -1, -1], np.int32), ["Water", "PMT", "Liner"], rows,
                    row_of, sensor_of_vol, names, groups, Params(values),
                    source_box=((-1.8, -1.8, -1.3), (1.8, 1.8, 1.3)))


def muon_track(seed: int = 2, length: float = 1.0, r_fid: float = 1.8, hz_fid: float = 1.3):
    """Return (start, direction) of a straight track of ``length`` m drawn uniformly in the fiducial cylinder with an
    isotropic direction, redrawn (seeded) until the whole track is inside it."""
    rng = np.random.default_rng(seed)
    while True:
        r = r_fid * np.sqrt(rng.uniform())
        ph = rng.uniform(0, 2 * np.pi)
        x0 = np.array([r * np.cos(ph), r * np.sin(ph), rng.uniform(-hz_fid, hz_fid)])
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        x1 = x0 + length * d
        if np.hypot(x1[0], x1[1]) < r_fid and abs(x1[2]) < hz_fid:
            return x0, d
