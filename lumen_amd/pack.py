"""Packing: a detector's parameter values into the flat operands the photon loop reads (the reference's
``build/pack.py:84-97`` ``pack.pack``, laid out for the HIP kernel and the C ABI).

Two arrays per run: ``values`` (float32) holds every table as its knots then its values, ``[E_0..E_{k-1},
V_0..V_{k-1}]``; ``layout`` (int32, fixed by the detector's structure) says where each table is.  A constant is a
table of one knot.  ``slots`` (int32, one per ``values`` entry) names the gradient slot of each table value the run
differentiates (-1: none).  Layout, in int32 words (``include/lumen_hip.h`` declares the same):

    [0] magic  [1] media  [2] classes  [3] rows  [4] sensors  [5] qe knots  [6] media at  [7] class_medium at
    [8] row_of at  [9] rows at  [10] qe table (offset of its knots in values)  [11] time nodes nt (0: no time
    histogram)  [12] t0 (ns, float32 bits)  [13] dt (ns, float32 bits)  [14] theta knots (1: no theta axis)
    [15] ce table (offset of its theta knots, radians, in values)  [16] exit-sensor flags at (C x C, 1: a detected
    share crossing that class pair is credited to the sensor of the volume left; 0: no table)  [17] born rows at (per
    class the row whose detected share a photon emitted inside a sensor volume of the class deposits at its emission,
    -1 none; 0: no table)
    The charge cells per sensor are (qe knots) x (theta knots), row-major (the wavelength's knot major).
    media:  per medium 25 words: (offset, knots) of RINDEX, of the group index, of ABSLENGTH, of RAYLEIGH; flags
            (1 opaque, 2 a wavelength shifter, 4 its exponential time profile, 8 Mie scattering, 16 ABSLENGTH and 32
            RAYLEIGH on RINDEX's knots); (offset, knots)
            of WLSABSLENGTH, of WLSCOMPONENT (emission intensity over energy), of WLSMEANNUMBERPHOTONS, of
            WLSTIMECONSTANT; of MIEHG, of MIEHG_FORWARD, of MIEHG_BACKWARD, of MIEHG_FORWARD_RATIO
    rows:   per row 18 words: model, flags, then (offset, knots) of R, T, eff, p_spike, p_lobe, p_back, sigma_alpha,
            polish

The group index is Geant4's GROUPVEL rule on the RINDEX knots (``media/laws.py:150-181``); a Rayleigh table
``"water_rule"`` is Geant4's Einstein-Smoluchowski table for Water on the RINDEX knots (``rules/optics.py:189-254``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import units
from .detector.model import COLUMNS, FLAG_EDGE, Detector
from .params import Params, Table

MAGIC = 0x4C484F31
HEADER = 18
MEDIUM_WORDS = 25
MED_OPAQUE, MED_WLS, MED_WLS_EXP, MED_MIE = 1, 2, 4, 8
MED_SAME_ABS, MED_SAME_RAY = 16, 32  # ABSLENGTH (RAYLEIGH) tabulated on RINDEX's knots
MIE_G_MAX = 0.999  # the largest |g| of a Mie lobe (the reference's rules/choices.py MIE_G_MAX)
ROW_WORDS = 2 + 2 * len(COLUMNS)


def group_index_knots(e, n):
    """Return Geant4's GROUPVEL knots as group indices (energies, n_g) from RINDEX knots (e, n).

    G4MaterialPropertiesTable::CalculateGROUPVEL as the reference restates it (``media/laws.py:150-166``): knot 0 at
    e_0 with the first segment's slope, knot j (1..k-1) at the midpoint of segment (j-1, j) with that segment's
    slope and mean index, knot k at e_k with the last segment's slope; n_g = max(n + dn/dlog E, n).
    """
    e, n = np.asarray(e, np.float64), np.asarray(n, np.float64)
    k = len(e) - 1
    if k < 1:
        return e.copy(), n.copy()
    eg = np.concatenate([e[:1], 0.5 * (e[:-2] + e[1:-1]), e[-1:]])
    ng = np.zeros(k + 1)
    for j in range(k + 1):
        s = min(max(j - 1, 0), k - 1)
        slope = (n[s + 1] - n[s]) / np.log(e[s + 1] / e[s])
        base = n[j] if j in (0, k) else 0.5 * (n[s] + n[s + 1])
        ng[j] = max(base + slope, base)
    return eg, ng


def water_rayleigh(e, n):
    """Return Geant4's Rayleigh lengths (m) of Water at RINDEX knots (e eV, n): 1/L = c1 k^4 ((n^2-1)(n^2+2)/3)^2."""
    e, n = np.asarray(e, np.float64), np.asarray(n, np.float64)
    c1 = units.WATER_COMPRESSIBILITY * units.WATER_TEMPERATURE * units.K_BOLTZMANN / (6.0 * np.pi)
    k = 2.0 * np.pi / units.HC_EV_M * e
    n2 = n * n
    return 1.0 / (c1 * k**4 * ((n2 - 1.0) * (n2 + 2.0) / 3.0) ** 2)


@dataclass
class Packed:
    """A run's packed operands: ``values`` float32, ``layout`` int32, ``slots`` int32, and the gradient slot map
    ``grad_keys`` = [(key, first slot, knots)]."""

    values: np.ndarray
    layout: np.ndarray
    slots: np.ndarray
    n_grad: int
    grad_keys: list


class _Builder:
    def __init__(self, grad_keys):
        self.vals, self.slots = [], []
        self.grad = {k: None for k in grad_keys}
        self.n_grad = 0
        self.grad_map = []

    def table(self, e, v, key=None):
        e, v = np.atleast_1d(np.asarray(e, np.float64)), np.atleast_1d(np.asarray(v, np.float64))
        if e.shape != v.shape or len(e) == 0:
            raise ValueError(f"pack: table {key!r} has {len(e)} knots and {len(v)} values")
        if len(e) > 1 and np.any(np.diff(e) <= 0):
            raise ValueError(f"pack: table {key!r}: knots must ascend")
        off = len(self.vals)
        self.vals += list(e) + list(v)
        slot = -1
        if key is not None and key in self.grad:
            if self.grad[key] is None:
                self.grad[key] = self.n_grad
                self.grad_map.append((key, self.n_grad, len(e)))
                self.n_grad += len(e)
            slot = self.grad[key]
        self.slots += [-1] * len(e) + ([slot + i for i in range(len(e))] if slot >= 0 else [-1] * len(e))
        return off, len(e)

    def value(self, p: Params, key):
        v = p[key]
        if isinstance(v, Table):
            return self.table(v.energy, v.value, key)
        return self.table([0.0] * np.size(v) if np.size(v) == 1 else np.arange(np.size(v), dtype=float),
                          np.atleast_1d(v), key)


def pack(det: Detector, params: Params | None = None, grad_keys=(), edge_rows="structure") -> Packed:
    """Return the Packed operands of ``det`` at ``params`` (default ``det.params()``).

    ``grad_keys``: the keys whose table values get gradient slots (the adjoint's parameters).  ``edge_rows``: the
    surface rows drawn with the defensive mixture (``transport/estimator.py:131-153``): "structure" (the default)
    the reference's rule, every row reading a keyed parameter that can zero an outcome (``SurfaceRow.edge``); or an
    explicit collection of row indices (empty: none).
    """
    p = det.params() if params is None else params
    if isinstance(edge_rows, str):
        edge_rows = {r for r, row in enumerate(det.rows) if row.edge()}
    for row in det.rows:  # a complex index sits in another column's slot: no derivative reads it
        bad = set(row.alias_keys()) & set(grad_keys)
        if bad:
            raise ValueError(f"pack: no derivative in {sorted(bad)} (surface {row.name!r}, {row.type}): the adjoint "
                             "has no terms for a complex index")
    for r, row in enumerate(det.rows):  # a GLISUR interface's Fresnel choices at each facet read both indices
        if row.type != "GLISUR_DIELECTRIC":
            continue
        for a_, b_ in zip(*np.nonzero(np.asarray(det.row_of) == r)):
            for c in (a_, b_):
                m = det.class_medium[c]
                if m >= 0 and det.media[m].rindex in grad_keys:
                    raise ValueError(f"pack: no derivative in {det.media[m].rindex!r}: surface {row.name!r} "
                                     "(GLISUR_DIELECTRIC) chooses at each facet by the indices, and the adjoint has "
                                     "no index terms for its facet path")
    b = _Builder(grad_keys)
    nm, nc, nr = len(det.media), len(det.class_medium), len(det.rows)
    med = np.zeros((nm, MEDIUM_WORDS), np.int32)
    for i, m in enumerate(det.media):
        if m.rindex is None:
            med[i, 8] = MED_OPAQUE
            continue
        rt = p[m.rindex]
        med[i, 0:2] = b.table(rt.energy, rt.value, m.rindex)
        eg, ng = group_index_knots(rt.energy, rt.value)
        med[i, 2:4] = b.table(eg, ng)
        if m.absorption is not None:
            med[i, 4:6] = b.value(p, m.absorption)
        if m.rayleigh == "water_rule":
            med[i, 6:8] = b.table(rt.energy, water_rayleigh(rt.energy, rt.value))
        elif m.rayleigh is not None:
            med[i, 6:8] = b.value(p, m.rayleigh)
        knots = np.float32(b.vals[med[i, 0]:med[i, 0] + med[i, 1]])
        for w, flag in ((4, MED_SAME_ABS), (6, MED_SAME_RAY)):  # on RINDEX's knots: one knot search serves both
            o, k = med[i, w], med[i, w + 1]
            if k == med[i, 1] and k > 1 and np.array_equal(np.float32(b.vals[o:o + k]), knots):
                med[i, 8] |= flag
        if m.wls_length is not None:
            med[i, 8] |= MED_WLS | (MED_WLS_EXP if m.wls_profile == "exponential" else 0)
            med[i, 9:11] = b.value(p, m.wls_length)
            spec = p[m.wls_spectrum]
            e, v = np.asarray(spec.energy, np.float64), np.asarray(spec.value, np.float64)
            if len(e) < 2 or np.any(v < 0) or not np.sum(0.5 * (v[1:] + v[:-1]) * np.diff(e)) > 0:
                raise ValueError(f"pack: {m.wls_spectrum}: an emission spectrum needs two or more knots and a "
                                 "non-negative intensity with positive integral")
            med[i, 11:13] = b.table(e, v, m.wls_spectrum)
            for j, c in ((13, m.wls_yield), (15, m.wls_tau)):
                med[i, j:j + 2] = b.value(p, c) if isinstance(c, str) else b.table([0.0], [float(c)])
        if m.mie_length is not None:
            med[i, 8] |= MED_MIE
            lt = p[m.mie_length]
            lv = np.atleast_1d(np.asarray(lt.value if isinstance(lt, Table) else lt, np.float64))
            if np.any(lv <= 0.0):
                raise ValueError(f"pack: {m.mie_length}: a scattering length of 0 (Geant4 scatters such a photon "
                                 "where it stands and never ends it)")
            med[i, 17:19] = b.value(p, m.mie_length)
            lobe = []
            for j, c in ((19, m.mie_forward), (21, m.mie_backward), (23, m.mie_ratio)):
                v = np.asarray(p[c] if isinstance(c, str) else c, np.float64)
                if v.size != 1:
                    raise ValueError(f"pack: medium {m.name!r}: a Mie lobe constant is one value, got {v.size}")
                lobe.append(float(v.reshape(-1)[0]))
                med[i, j:j + 2] = b.table([0.0], [lobe[-1]], c if isinstance(c, str) else None)
            if abs(lobe[0]) > MIE_G_MAX or abs(lobe[1]) > MIE_G_MAX or not 0.0 <= lobe[2] <= 1.0:
                raise ValueError(f"pack: medium {m.name!r}: Mie lobes need |g_f|, |g_b| <= {MIE_G_MAX} and a forward "
                                 f"ratio in [0, 1], got {lobe}")
    rows = np.zeros((nr, ROW_WORDS), np.int32)
    for r, row in enumerate(det.rows):
        rows[r, 0] = row.model
        rows[r, 1] = row.flags() | (FLAG_EDGE if r in edge_rows else 0)
        for j, c in enumerate(COLUMNS):
            v = row.column(c)
            rows[r, 2 + 2 * j: 4 + 2 * j] = b.value(p, v) if isinstance(v, str) else b.table([0.0], [float(v)])
        if row.type.startswith("GLISUR") and isinstance(row.columns["polish"], str):
            pv = p[row.columns["polish"]]
            for x in np.atleast_1d(pv.value if isinstance(pv, Table) else pv):
                row.check_polish(float(x))
    qe_off, n_qe = 0, 1
    q = p[det.groups[0].qe] if det.groups and det.groups[0].qe is not None else None
    if isinstance(q, Table):  # the charge histogram's wavelength knots (nm); QE itself is read after tracking
        qe_off, n_qe = b.table(q.energy, q.value)
    else:  # no QE table, or one QE per sensor: a single charge cell per sensor
        qe_off, n_qe = b.table([0.0], [1.0])
    ce = p[det.groups[0].ce] if det.groups and getattr(det.groups[0], "ce", None) is not None else None
    if isinstance(ce, Table) and len(ce.energy) > 1:  # the charge histogram's theta knots (rad); CE read after tracking
        ce_off, n_ce = b.table(ce.energy, ce.value)
    else:
        ce_off, n_ce = b.table([0.0], [1.0])
    tg = getattr(det, "time_grid", None)
    f32bits = lambda x: int(np.array([x], np.float32).view(np.int32)[0])  # noqa: E731
    L = [MAGIC, nm, nc, nr, det.n_sensors, n_qe, 0, 0, 0, 0, qe_off,
         0 if tg is None else tg.nt, 0 if tg is None else f32bits(tg.t0), 0 if tg is None else f32bits(tg.dt),
         n_ce, ce_off, 0, 0]
    L[6] = len(L)
    L += list(med.ravel())
    L[7] = len(L)
    L += list(np.asarray(det.class_medium, np.int32))
    L[8] = len(L)
    L += list(np.asarray(det.row_of, np.int32).ravel())
    L[9] = len(L)
    L += list(rows.ravel())
    left = getattr(det, "credit_left", None)
    if left is not None and np.any(left):
        L[16] = len(L)
        L += list(np.asarray(left, np.int32).ravel())
    born = getattr(det, "born_row", None)
    if born is not None and np.any(np.asarray(born) >= 0):
        L[17] = len(L)
        L += list(np.asarray(born, np.int32))
    return Packed(np.asarray(b.vals, np.float32), np.asarray(L, np.int32), np.asarray(b.slots, np.int32),
                  b.n_grad, b.grad_map)
