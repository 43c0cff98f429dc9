"""The liquid-argon TPC-like box: a 5.0 x 4.0 x 5.0 m LAr volume with PTFE-lined walls and 120 PMTs behind
wavelength-shifting PMMA discs on one wall, built directly as navigator arrays and optics (no GDML).

- LAr: RINDEX 1.38 at 128 nm falling to 1.23 at 420 nm (knots in photon energy), Rayleigh length 0.90 m at 128 nm
  scaling as lambda^4, absorption length 20 m.
- Walls: the border LAr -> outside is a PTFE-like reflector, Geant4's groundfrontpainted (``PAINTED`` with f_spec 0:
  reflectivity 0.95, Lambertian).
- 120 PMTs on the y = -2 m wall (a 12 x 10 grid, pitch 5/12 m along x and 0.5 m along z): each a cylinder of radius
  0.2 m and 50 mm depth (its cathode the border LAr -> PMT, R = 0, EFFICIENCY 1, QE 0.25 after tracking), and 5 mm
  in front of it a 3 mm PMMA disc of radius 0.2 m whose wavelength-shifting coating is a bulk WLS channel (G4OpWLS):
  WLSABSLENGTH 0.1 mm at VUV energies and 1 km in the visible, emission peaked at 420 nm, WLSMEANNUMBERPHOTONS 0.5 (the
  conversion efficiency), immediate (delta profile).  LAr <-> PMMA is a polished dielectric interface (Fresnel).
- Sources: 4,096 isotropic 128 nm points at the centres of a 16 x 16 x 16 voxel grid of the box.

No reference form exists for a converting boundary, so the coating is modelled, as Geant4 applications do, as a
thin bulk wavelength shifter (the reference's ``media/bulk.py`` channel "wls" with ``media/emission.py``'s record).
"""

from __future__ import annotations

import numpy as np

from .. import units
from ..params import Params, Table
from . import navarrays, solids
from .model import Detector, Medium, SensorGroup, SurfaceRow

BOX_H = (2.5, 2.0, 2.5)  # half lengths (m)
PMT_R, PMT_HALF = 0.2, 0.025
DISC_HALF, DISC_GAP, WALL_GAP = 0.0015, 0.005, 0.003
N_X, N_Z = 12, 10
LAM_VUV, LAM_VIS = 128.0, 420.0


def _lam_table(lams, vals):
    """Return a Table over photon energy (ascending) from values at wavelengths ``lams`` (nm)."""
    e = units.HC / np.asarray(lams, np.float64)
    order = np.argsort(e)
    return Table(e[order], np.asarray(vals, np.float64)[order])


def lar_tables():
    """Return (RINDEX, ABSLENGTH, RAYLEIGH) of LAr on knots from 110 to 650 nm."""
    lam = np.array([110.0, 118.0, 128.0, 140.0, 160.0, 200.0, 260.0, 340.0, 420.0, 520.0, 650.0])
    # n = 1.23 + 0.15 (x - x_420) / (1 - x_420) with x = (128 nm / lambda)^2: 1.38 at 128 nm, 1.23 at 420 nm
    x, x420 = (LAM_VUV / lam) ** 2, (LAM_VUV / LAM_VIS) ** 2
    n = 1.23 + 0.15 * (x - x420) / (1.0 - x420)
    ray = 0.90 * (lam / LAM_VUV) ** 4
    return _lam_table(lam, n), _lam_table(lam, np.full_like(lam, 20.0)), _lam_table(lam, ray)


def pmma_tables():
    """Return (RINDEX, ABSLENGTH, WLSABSLENGTH, WLSCOMPONENT) of the coated PMMA disc."""
    lam = np.array([110.0, 128.0, 200.0, 250.0, 300.0, 350.0, 380.0, 420.0, 500.0, 650.0])
    n = np.interp(lam, [110.0, 128.0, 420.0, 650.0], [1.62, 1.60, 1.50, 1.49])
    wls = np.where(lam <= 250.0, 1e-4, np.where(lam <= 300.0, 1e-3, 1000.0))
    spec_lam = np.array([380.0, 400.0, 420.0, 440.0, 460.0])
    spec = np.array([0.0, 0.5, 1.0, 0.5, 0.0])
    return (_lam_table(lam, n), _lam_table(lam, np.full_like(lam, 10.0)), _lam_table(lam, wls),
            _lam_table(spec_lam, spec))


def _y_frame():
    """Return the rotation taking a tube's axis (local z) to world +y."""
    return np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, -1.0, 0.0]]).T


def pmt_centres():
    """Return the (x, z) centres of the 120 PMTs on the y = -2 m wall."""
    hx, _, hz = BOX_H
    xs = -hx + (np.arange(N_X) + 0.5) * (2 * hx / N_X)
    zs = -hz + (np.arange(N_Z) + 0.5) * (2 * hz / N_Z)
    return np.array([[x, z] for z in zs for x in xs])


def lar_box(reflectivity=0.95, qe=0.25, yield_=0.5) -> Detector:
    """Return the LAr box (module docstring)."""
    hx, hy, hz = BOX_H
    kinds = [solids.box(hx, hy, hz, name="lar"), solids.tube(PMT_R, DISC_HALF, name="disc"),
             solids.tube(PMT_R, PMT_HALF, name="pmt")]
    R = _y_frame()
    vols = [navarrays.Volume("lar", 0, np.eye(3), np.zeros(3), -1, 0)]
    centres = pmt_centres()
    y_pmt = -hy + WALL_GAP + PMT_HALF
    y_disc = -hy + WALL_GAP + 2 * PMT_HALF + DISC_GAP + DISC_HALF
    for i, (x, z) in enumerate(centres):
        vols.append(navarrays.Volume(f"lar/disc_{i}", 1, R, np.array([x, y_disc, z]), 0, 1))
    for i, (x, z) in enumerate(centres):
        vols.append(navarrays.Volume(f"lar/pmt_{i}", 2, R, np.array([x, y_pmt, z]), 0, 2))
    nav = navarrays.build(vols, kinds, outside_cls=3)
    n_lar, a_lar, r_lar = lar_tables()
    n_pm, a_pm, wls_len, wls_spec = pmma_tables()
    n_pmt = len(centres)
    values = {
        "materials/LAr/RINDEX": n_lar,
        "materials/LAr/ABSLENGTH": a_lar,
        "materials/LAr/RAYLEIGH": r_lar,
        "materials/PMMA/RINDEX": n_pm,
        "materials/PMMA/ABSLENGTH": a_pm,
        "materials/PMMA/WLSABSLENGTH": wls_len,
        "materials/PMMA/WLSCOMPONENT": wls_spec,
        "materials/PMMA/WLSMEANNUMBERPHOTONS": np.float64(yield_),
        "surfaces/PTFE/REFLECTIVITY": np.float64(reflectivity),
        "surfaces/Cathode/REFLECTIVITY": np.float64(0.0),
        "surfaces/Cathode/EFFICIENCY": np.float64(1.0),
        "sensors/pmt/qe": np.full(n_pmt, float(qe)),
        "sensors/pmt/efficiency": np.float64(1.0),
    }
    media = [Medium("LAr", "materials/LAr/RINDEX", "materials/LAr/ABSLENGTH", "materials/LAr/RAYLEIGH"),
             Medium("PMMA", "materials/PMMA/RINDEX", "materials/PMMA/ABSLENGTH",
                    wls_length="materials/PMMA/WLSABSLENGTH", wls_spectrum="materials/PMMA/WLSCOMPONENT",
                    wls_yield="materials/PMMA/WLSMEANNUMBERPHOTONS")]
    rows = [SurfaceRow("PTFE", "PAINTED", {"R": "surfaces/PTFE/REFLECTIVITY", "p_spike": 0.0}),
            SurfaceRow("LAr_PMMA", "DIELECTRIC"),
            SurfaceRow("Cathode", "SPECULAR_METAL", {"R": "surfaces/Cathode/REFLECTIVITY",
                                                     "eff": "surfaces/Cathode/EFFICIENCY"})]
    row_of = np.full((4, 4), -1, np.int32)
    row_of[0, 3] = 0  # LAr -> outside: the PTFE foil
    row_of[0, 1] = 1  # LAr <-> PMMA: Fresnel
    row_of[1, 0] = 1
    row_of[0, 2] = 2  # LAr -> PMT: the cathode
    sensor_of_vol = np.full(nav.n + 1, -1, np.int32)
    first_pmt = 1 + n_pmt
    sensor_of_vol[first_pmt: first_pmt + n_pmt] = np.arange(n_pmt)
    names = [v.name for v in vols[first_pmt:]]
    groups = [SensorGroup("pmt", np.arange(n_pmt), "sensors/pmt/qe", "sensors/pmt/efficiency")]
    return Detector("lar_box_120", nav, media, np.array([0, 1, -1, -1], np.int32), ["LAr", "PMMA", "PMT", "PTFE"],
                    rows, row_of, sensor_of_vol, names, groups, Params(values),
                    source_box=((-hx, -hy, -hz), (hx, hy, hz)))


def voxel_points(n=16, lam=LAM_VUV, photons=float(1 << 16)):
    """Return the sources at the centres of an n x n x n voxel grid of the box: isotropic ``lam`` nm points."""
    from ..sources import Point

    hx, hy, hz = BOX_H
    c = [(-h + (np.arange(n) + 0.5) * (2 * h / n)) for h in (hx, hy, hz)]
    return [Point((float(x), float(y), float(z)), lam=lam, n_photons=photons)
            for x in c[0] for y in c[1] for z in c[2]]
