"""A detector as the photon loop reads it: navigator arrays, media, classes, the surface table and the sensors.

This restates the structure of the reference's built detector on the path (``build/lookup.py:360-390`` the surface
table, ``transport/step.py:280-295`` rows by class pair, ``media/bulk.py:76-110`` the media, ``surfaces/types.py``
the surface types) for the types the engine implements:

====================  =======  ========================================================================
type                   model    reads (columns)
====================  =======  ========================================================================
ABSORBER               reflect  nothing: every photon absorbed, nothing detected
SPECULAR_METAL         reflect  R (mirror), eff (the absorbed share detected)
UNIFIED_METAL          reflect  R split by p_spike / p_lobe (facet path, sigma_alpha) / p_back / Lambertian, eff
UNIFIED_METAL_NORMAL   reflect  as UNIFIED_METAL at sigma_alpha 0: the lobe reflects about the global normal
PAINTED                reflect  R split by f_spec (p_spike column: mirror) and Lambertian, eff
PAINTED_T              reflect  PAINTED with T (straight on): Geant4's frontpainted finish with TRANSMITTANCE
COUNTING               reflect  eff: every photon absorbed, the share eff detected (a counting-only sensor)
DIELECTRIC             fresnel  the Fresnel interface of the two media (R fixed at 1)
DIELECTRIC_RT          fresnel  R (into the interface), T (straight on), eff
PHOTOCATHODE           complex  nr, nk (a complex index seen from the dielectric: its Fresnel reflectance is the
                                mirror's share, the absorbed rest detected with eff); nr, nk in the p_lobe and
                                p_back column slots, which no other part of the type reads
GLISUR_METAL           reflect  R into GLISUR's facet path (radius 1 - polish), eff: Geant4's dielectric_metal,
                                glisur, ground, polish < 1
GLISUR_DIELECTRIC      reflect  R into GLISUR's facet path that reflects or refracts at each facet by the Fresnel
                                shares of the two media (it may end beyond the surface), T straight on, eff:
                                Geant4's dielectric_dielectric, glisur, ground, polish < 1
====================  =======  ========================================================================

A row with ``transmitting`` set (Geant4's dielectric_metal with TRANSMITTANCE, the reference's
``surfaces/types.py:497-512`` ``transmitting``) also reads T: the share past R up to R + T goes straight on.

A GLISUR row whose polish is a parameter (a key) draws its facets from the cover mixture (1 - e) p_m + e p_shell
(the reference's ``surfaces/facets.py:14-27, 557-583``): polish moves the facet support's edge.

``row_of[c_from, c_to]`` is the surface row of a crossing from class ``c_from`` into ``c_to`` (-1: no surface
joins them; the photon is lost).  A volume's class decides its medium (``class_medium``).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ..params import Params, Table
from .navarrays import NavArrays

# surface models (the kernel's branches) and the six continuing outcomes, as the reference's
# ``surfaces/types.py:288-296`` and ``surfaces/outcomes.py:36-53``
MODEL_REFLECT, MODEL_FRESNEL, MODEL_COMPLEX = 0, 1, 2
STRAIGHT, REFRACT, SPECULAR, BACKSCATTER, GLOSSY, DIFFUSE = range(6)
COLUMNS = ("R", "T", "eff", "p_spike", "p_lobe", "p_back", "sigma_alpha", "polish")
NEUTRAL = dict(R=0.0, T=0.0, eff=0.0, p_spike=0.0, p_lobe=0.0, p_back=0.0, sigma_alpha=0.1, polish=0.0)
FLAG_NORMAL, FLAG_FAR_OPAQUE, FLAG_EDGE, FLAG_PER_NORMAL = 1, 2, 4, 8
FLAG_GLISUR, FLAG_REFRACTS, FLAG_COVER = 16, 32, 64  # GLOSSY is GLISUR's path; it refracts; drawn under the cover
GLISUR_R_MIN = 1e-3  # the smallest 1 - polish of a GLISUR facet path (rules/choices.py:97-100)

# Geant4 property names -> columns
G4_COLUMN = {"REFLECTIVITY": "R", "TRANSMITTANCE": "T", "EFFICIENCY": "eff", "SPECULARSPIKECONSTANT": "p_spike",
             "SPECULARLOBECONSTANT": "p_lobe", "BACKSCATTERCONSTANT": "p_back", "sigma_alpha": "sigma_alpha",
             "REALRINDEX": "nr", "IMAGINARYRINDEX": "nk"}

# the outcomes each type can give with its constants fixed (its support: what the defensive mixture covers,
# ``surfaces/types.py:404-440`` support rules), as bits by outcome index
SUPPORT = {
    "ABSORBER": (),
    "SPECULAR_METAL": (SPECULAR,),
    "UNIFIED_METAL": (SPECULAR, BACKSCATTER, GLOSSY, DIFFUSE),
    "UNIFIED_METAL_NORMAL": (SPECULAR, BACKSCATTER, DIFFUSE),
    "PAINTED": (SPECULAR, DIFFUSE),
    "PAINTED_T": (STRAIGHT, SPECULAR, DIFFUSE),
    "COUNTING": (),
    "DIELECTRIC": (REFRACT, SPECULAR),
    "DIELECTRIC_RT": (REFRACT, SPECULAR, STRAIGHT),
    "PHOTOCATHODE": (SPECULAR,),
    "GLISUR_METAL": (GLOSSY,),
    "GLISUR_DIELECTRIC": (GLOSSY, STRAIGHT),
}

# type: (model, the parameters it reads, its constants, its flags, {parameter: the column slot it is read from})
TYPES = {
    "ABSORBER": (MODEL_REFLECT, (), {}, 0, {}),
    "SPECULAR_METAL": (MODEL_REFLECT, ("R", "eff"), {"p_spike": 1.0}, FLAG_NORMAL, {}),
    "UNIFIED_METAL": (MODEL_REFLECT, ("R", "p_spike", "p_lobe", "p_back", "sigma_alpha", "eff"), {}, 0, {}),
    "UNIFIED_METAL_NORMAL": (MODEL_REFLECT, ("R", "p_spike", "p_lobe", "p_back", "eff"), {}, FLAG_NORMAL, {}),
    "PAINTED": (MODEL_REFLECT, ("R", "p_spike", "eff"), {}, 0, {}),
    "PAINTED_T": (MODEL_REFLECT, ("R", "T", "p_spike", "eff"), {}, 0, {}),
    "COUNTING": (MODEL_REFLECT, ("eff",), {}, 0, {}),
    "DIELECTRIC": (MODEL_FRESNEL, (), {"R": 1.0}, 0, {}),
    "DIELECTRIC_RT": (MODEL_FRESNEL, ("R", "T", "eff"), {}, 0, {}),
    "PHOTOCATHODE": (MODEL_COMPLEX, ("nr", "nk", "eff"), {}, 0, {"nr": "p_lobe", "nk": "p_back"}),
    "GLISUR_METAL": (MODEL_REFLECT, ("R", "polish", "eff"), {"p_lobe": 1.0}, FLAG_GLISUR, {}),
    "GLISUR_DIELECTRIC": (MODEL_REFLECT, ("R", "T", "polish", "eff"), {"p_lobe": 1.0}, FLAG_GLISUR | FLAG_REFRACTS,
                          {}),
}
ZERO_ROLE = ("R", "T", "eff", "p_spike", "p_lobe", "p_back")  # parameters that can zero an outcome's probability


WLS_PROFILES = ("delta", "exponential")  # G4OpWLS's time profiles (delta: Geant4's default)


@dataclass
class Medium:
    """A bulk medium: its keys (None: absent).  ``rayleigh`` may be ``"water_rule"``: Geant4's table from RINDEX.

    A wavelength shifter (G4OpWLS, the reference's ``media/bulk.py`` channel "wls" with ``media/emission.py``'s
    record): ``wls_length`` the WLSABSLENGTH key; ``wls_spectrum`` the WLSCOMPONENT key (a Table of emission
    intensity over photon energy); ``wls_yield`` WLSMEANNUMBERPHOTONS (a key or a constant; 1 when absent: the
    expected photons per absorbed one, a weight); ``wls_tau`` WLSTIMECONSTANT (ns; a key or a constant, 0 when
    absent) under the time profile ``wls_profile``.  The emission is isotropic and unpolarised, at or below the
    absorbed energy.

    Mie scattering (G4OpMieHG, ``media/bulk.py`` ``mie``, ``media/phase.py:98-190``): ``mie_length`` the MIEHG key
    (a mean free path over photon energy); ``mie_forward``, ``mie_backward``, ``mie_ratio`` MIEHG_FORWARD,
    MIEHG_BACKWARD, MIEHG_FORWARD_RATIO (keys or constants): the forward lobe HG(g_f) with probability r, else the
    backward one HG(-g_b); |g| <= 0.999, 0 <= r <= 1.
    """

    name: str
    rindex: str | None
    absorption: str | None = None
    rayleigh: str | None = None
    mie_length: str | None = None
    mie_forward: str | float = 0.0
    mie_backward: str | float = 0.0
    mie_ratio: str | float = 1.0
    wls_length: str | None = None
    wls_spectrum: str | None = None
    wls_yield: str | float = 1.0
    wls_tau: str | float = 0.0
    wls_profile: str = "delta"

    def __post_init__(self):
        if self.wls_length is not None and self.wls_spectrum is None:
            raise ValueError(f"medium {self.name!r}: WLSABSLENGTH without WLSCOMPONENT (G4OpWLS absorbs and "
                             "emits nothing)")
        if self.wls_profile not in WLS_PROFILES:
            raise ValueError(f"medium {self.name!r}: wls_profile {self.wls_profile!r} is not one of {WLS_PROFILES}")


@dataclass
class SurfaceRow:
    """One row of the surface table: its type, and per column a parameter key or a constant."""

    name: str
    type: str
    columns: dict = field(default_factory=dict)  # parameter -> key (str) or number
    far_opaque: bool = False
    transmitting: bool = False  # also reads T: Geant4's dielectric_metal with TRANSMITTANCE
    per_normal: bool = False  # read for a QE applied after tracking: the detected share over A_n (types.py:576)

    def __post_init__(self):
        if self.type not in TYPES:
            raise ValueError(f"surface row {self.name!r}: unknown type {self.type!r} (one of {sorted(TYPES)})")
        if self.transmitting and self.type not in ("SPECULAR_METAL", "UNIFIED_METAL", "UNIFIED_METAL_NORMAL",
                                                   "PHOTOCATHODE", "GLISUR_METAL"):
            raise ValueError(f"surface row {self.name!r}: only a dielectric_metal type transmits (T)")
        if self.per_normal and "eff" not in TYPES[self.type][1]:
            raise ValueError(f"surface row {self.name!r}: per_normal needs a detecting type (one reading eff)")
        if self.far_opaque and TYPES[self.type][3] & FLAG_REFRACTS:
            raise ValueError(f"surface row {self.name!r}: {self.type}'s facet path decides its own side (no "
                             "far_opaque)")
        if self.type.startswith("GLISUR"):
            if "polish" not in self.columns:
                raise ValueError(f"surface row {self.name!r}: {self.type} reads polish")
            self.check_polish(self._fixed("polish"))

    def check_polish(self, v):
        """Refuse a GLISUR polish outside [0, 1 - GLISUR_R_MIN] (None: a key, checked with its value)."""
        if v is not None and not (0.0 <= v <= 1.0 - GLISUR_R_MIN):
            raise ValueError(f"surface row {self.name!r}: polish {v} outside [0, {1 - GLISUR_R_MIN}] (polish 1 is "
                             "the polished reading: SPECULAR_METAL or DIELECTRIC_RT)")

    @property
    def model(self) -> int:
        return TYPES[self.type][0]

    def params(self) -> tuple:
        """The parameters the row reads: its type's, and T when it transmits."""
        p = TYPES[self.type][1]
        return p + ("T",) if self.transmitting and "T" not in p else p

    def _fixed(self, name):
        """A parameter's value when the row holds it as a constant (not a key), else None."""
        v = self.columns.get(name)
        if isinstance(v, str):
            return None
        if v is None:  # no key: the column holds the type's constant, else the neutral value
            return float(TYPES[self.type][2].get(name, NEUTRAL.get(name, 0.0)))
        return float(v)

    def flags(self) -> int:
        """Return the row's flag word: type flags, far-opaque, and its support's outcome bits at bit 8: the
        outcomes its fixed constants leave possible (the reference's support rules, ``surfaces/types.py:404-440``
        and ``transmitting``)."""  # noqa: D205, D209
        sup = set(SUPPORT[self.type])
        if self.type in ("PAINTED", "PAINTED_T"):  # a fixed f_spec of 0 (1) leaves no mirror (Lambertian)
            fs = self._fixed("p_spike")
            sup -= {SPECULAR} if fs == 0.0 else {DIFFUSE} if fs == 1.0 else set()
        if self.type == "UNIFIED_METAL_NORMAL":
            ps, pl, pb = self._fixed("p_spike"), self._fixed("p_lobe"), self._fixed("p_back")
            if ps == 0.0 and pl == 0.0:
                sup.discard(SPECULAR)
            if pb == 0.0:
                sup.discard(BACKSCATTER)
            if None not in (ps, pl, pb) and ps + pl + pb >= 1.0:
                sup.discard(DIFFUSE)
        if self.transmitting or self.type in ("PAINTED_T", "GLISUR_DIELECTRIC"):
            if self._fixed("T") == 0.0:
                sup.discard(STRAIGHT)
            elif self.transmitting:
                sup.add(STRAIGHT)
        if self.far_opaque:
            sup -= {STRAIGHT, REFRACT}
        bits = sum(1 << k for k in sup)
        cover = FLAG_COVER if TYPES[self.type][3] & FLAG_GLISUR and isinstance(self.columns.get("polish"), str) else 0
        return (TYPES[self.type][3] | (FLAG_FAR_OPAQUE if self.far_opaque else 0) |
                (FLAG_PER_NORMAL if self.per_normal else 0) | cover | (bits << 8))

    def edge(self) -> bool:
        """Return whether the row is an edge row: a parameter it reads is keyed and can zero an outcome's
        probability (the cuts and the reflection split: the defensive role "zero", ``surfaces/types.py:322-331``;
        decided by structure, never by a value, as ``api.py:797-814`` ``_edge_rows``)."""  # noqa: D205, D209
        return any(isinstance(self.columns.get(c), str) for c in ZERO_ROLE
                   if c in self.params() and c not in TYPES[self.type][4])

    def column(self, c):
        """Return what column slot ``c`` holds for this row: the key or constant of the parameter read from it (an
        alias slot, e.g. a photocathode's nr in p_lobe), else the type's fixed value, else the neutral value."""  # noqa: D205, D209
        for name, slot in TYPES[self.type][4].items():
            if slot == c:
                return self.columns[name]
        if c in self.columns and c in self.params():
            return self.columns[c]
        return TYPES[self.type][2].get(c, NEUTRAL[c])

    def alias_keys(self) -> tuple:
        """The keys this row reads through an alias slot (no derivative: the slot's own rules do not apply)."""
        keys = [self.columns[n] for n in TYPES[self.type][4] if isinstance(self.columns.get(n), str)]
        if self.per_normal and self.model == MODEL_COMPLEX and isinstance(self.columns.get("T"), str):
            keys.append(self.columns["T"])  # a / A_n then depends on the crossing: no derivative term for T
        return tuple(keys)


@dataclass(frozen=True)
class TimeGrid:
    """The time histogram's nodes t0 + j dt (ns), j < nt: each deposit split over its four nearest nodes with cubic
    Lagrange weights; what falls beyond the last node (late) or before the first (early) kept as (q, q t) per sensor
    (the reference's ``readout/grid.py`` ``Grid`` t0, dt, nt and ``time_cells``)."""  # noqa: D205, D209

    t0: float
    dt: float
    nt: int

    def __post_init__(self):
        if self.nt < 1 or not self.dt > 0:
            raise ValueError(f"TimeGrid: nt >= 1 and dt > 0, got nt={self.nt}, dt={self.dt}")


@dataclass
class SensorGroup:
    """A sensor group: its sensors (indices into the detector's sensors), QE, efficiency and collection-efficiency
    keys.  ``ce``: a Table of the collection efficiency over the cathode position angle theta (rad: the polar angle
    of the crossing point about the sensor volume's +z axis, WCSim's angle; the reference's ``sensors/features.py``
    ``theta``): the charge histogram then gains a theta axis on its knots (hat weights) and ``readout.pe`` multiplies
    CE(theta) in, as QE(lambda)."""  # noqa: D205, D209

    name: str
    sensors: np.ndarray
    qe: str | None = None
    efficiency: str | None = None
    ce: str | None = None


@dataclass
class Detector:
    """A loaded detector (module docstring).  ``params()`` gives its parameter values (the drop-in ``det.params()``)."""

    name: str
    nav: NavArrays
    media: list
    class_medium: np.ndarray  # (C,) medium index per class (-1: opaque, no medium)
    class_names: list
    rows: list
    row_of: np.ndarray  # (C, C) int32
    sensor_of_vol: np.ndarray  # (n + 1,) int32, -1 none
    sensor_names: list
    groups: list
    values: Params
    source_box: tuple = None
    time_grid: "TimeGrid | None" = None  # the time histogram's static nodes (None: no time histogram)
    # the sensor-crediting rules (the reference's ``build/lookup.py:413-417`` ``deposit_pre``, ``born_pair``):
    # credit_left (C, C) bool, a detected share crossing c_from -> c_to goes to the sensor of the volume left (an
    # exit sensor: a border out of a sensor volume whose surface detects); born_row (C,) int32, the row whose detected
    # share a photon emitted inside a sensor volume of the class deposits at its emission, where it ends (-1: none)
    credit_left: np.ndarray | None = None
    born_row: np.ndarray | None = None

    def timed(self, t0: float, dt: float, nt: int) -> "Detector":
        """Return this detector with a time histogram on the nodes t0 + j dt, j < nt (ns)."""
        import dataclasses

        return dataclasses.replace(self, time_grid=TimeGrid(float(t0), float(dt), int(nt)))

    def params(self) -> Params:
        """Return the detector's parameter values."""
        return self.values

    @property
    def n_sensors(self) -> int:
        return len(self.sensor_names)

    @property
    def reach(self) -> float:
        """The farthest a straight path goes inside the geometry: the diagonal of the box holding every volume (m)."""
        lo, hi = self.nav.lo_w.min(0), self.nav.hi_w.max(0)
        return float(np.linalg.norm(hi - lo))

    def sensor_table(self):
        """Return (names, world positions (S, 3)) of the sensors in ``Light.arrivals`` order."""
        pos = np.zeros((self.n_sensors, 3))
        for v in range(self.nav.n):
            s = self.sensor_of_vol[v]
            if s >= 0:
                pos[s] = self.nav.host["t"][v]
        return list(self.sensor_names), pos

    def __repr__(self):
        return (f"{self.name}: {self.nav.n} volumes, {len(self.media)} media, {len(self.rows)} surface rows, "
                f"{self.n_sensors} sensors")


def table_at(tab: Table, x):
    """Return a table at x (its knot coordinate), linear between knots, held beyond (G4PhysicsVector::Value)."""
    return np.interp(x, np.asarray(tab.energy, np.float64), np.asarray(tab.value, np.float64))
