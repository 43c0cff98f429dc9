"""Toy detectors with closed-form light, built with the engine's own builder (the reference's physics toys,
tests/physics/test_toys.py:126-182): a water sphere of radius ``a`` inside a sensor shell, the crossing out of the
water a surface of one type, a point source at the centre.
"""

from __future__ import annotations

import numpy as np

from lumen_amd import units
from lumen_amd.detector import navarrays, solids
from lumen_amd.detector.model import Detector, Medium, SensorGroup, SurfaceRow
from lumen_amd.params import Params, Table

E_KNOTS = np.array([1.5, 5.0])  # eV: flat tables over every wavelength used here


WLS_KNOTS = np.array([1.5, 2.95, 3.05, 8.0])  # eV: a shifter absorbing above 3 eV only
SPEC_E = np.array([2.2, 2.4, 2.6, 2.8])  # eV: its emission (443-564 nm), below its absorption edge
SPEC_S = np.array([0.0, 1.0, 0.6, 0.0])


def sphere_toy(a=1.0, surface="SPECULAR_METAL", R=0.5, L_abs=10.0, L_ray=None, n=1.33, eff=1.0, L_wls=None,
               wls_yield=0.7, wls_tau=0.0, wls_profile="delta", ce=None, extra=None, transmitting=False,
               far_opaque=False, per_normal=False):
    """Return a water sphere (radius a) inside a sensor shell: the crossing out of the water is ``surface``
    (SPECULAR_METAL: R mirror; PAINTED: R Lambertian) detecting the absorbed share with ``eff``.  With ``L_wls``
    the water holds a wavelength shifter absorbing above 3 eV with mean free path L_wls (m) and emitting on
    2.2-2.8 eV with yield ``wls_yield``.  ``ce``: a Table of the shell's collection efficiency over the cathode
    position angle (rad, the polar angle of the crossing point about +z)."""  # noqa: D205, D209
    kinds = [solids.orb(a + 0.05, name="shell"), solids.orb(a, name="water")]
    vols = [navarrays.Volume("shell", 0, np.eye(3), np.zeros(3), -1, 1),
            navarrays.Volume("shell/water", 1, np.eye(3), np.zeros(3), 0, 0)]
    nav = navarrays.build(vols, kinds, outside_cls=2)
    values = {"materials/Water/RINDEX": Table(E_KNOTS, np.array([n, n]))}
    absorption = rayleigh = None
    if L_abs is not None:
        values["materials/Water/ABSLENGTH"] = Table(E_KNOTS, np.array([L_abs, L_abs]))
        absorption = "materials/Water/ABSLENGTH"
    if L_ray is not None:
        values["materials/Water/RAYLEIGH"] = Table(E_KNOTS, np.array([L_ray, L_ray]))
        rayleigh = "materials/Water/RAYLEIGH"
    wls = {}
    if L_wls is not None:
        values["materials/Water/WLSABSLENGTH"] = Table(WLS_KNOTS, np.array([1e30, 1e30, L_wls, L_wls]))
        values["materials/Water/WLSCOMPONENT"] = Table(SPEC_E, SPEC_S)
        values["materials/Water/WLSMEANNUMBERPHOTONS"] = np.float64(wls_yield)
        values["materials/Water/WLSTIMECONSTANT"] = np.float64(wls_tau)
        wls = dict(wls_length="materials/Water/WLSABSLENGTH", wls_spectrum="materials/Water/WLSCOMPONENT",
                   wls_yield="materials/Water/WLSMEANNUMBERPHOTONS", wls_tau="materials/Water/WLSTIMECONSTANT",
                   wls_profile=wls_profile)
    values["surfaces/Wall/REFLECTIVITY"] = np.float64(R)
    values["surfaces/Wall/EFFICIENCY"] = np.float64(eff)
    cols = {"R": "surfaces/Wall/REFLECTIVITY", "eff": "surfaces/Wall/EFFICIENCY"}
    if surface in ("PAINTED", "PAINTED_T"):
        cols["p_spike"] = 0.0
    for c, v in (extra or {}).items():  # more columns: a key (with its value) per name
        key = f"surfaces/Wall/{c}"
        values[key] = v if isinstance(v, Table) else np.float64(v)
        cols[c] = key
    rows = [SurfaceRow("Wall", surface, cols, far_opaque=far_opaque, transmitting=transmitting,
                       per_normal=per_normal),
            SurfaceRow("absorber", "ABSORBER")]
    row_of = np.full((3, 3), -1, np.int32)
    row_of[0, 1] = 0
    row_of[0, 2] = 1
    sensor_of_vol = np.array([0, -1, -1], np.int32)
    ce_key = None
    if ce is not None:
        values["sensors/pmt/ce"] = ce
        ce_key = "sensors/pmt/ce"
    return Detector("sphere_toy", nav, [Medium("Water", "materials/Water/RINDEX", absorption, rayleigh, **wls)],
                    np.array([0, -1, -1], np.int32), ["Water", "Shell", "OUT"], rows, row_of, sensor_of_vol,
                    ["shell"], [SensorGroup("pmt", np.arange(1), ce=ce_key)], Params(values))


def sensor_volume_toy(rule="born", eff=0.6, R=0.3, T=0.4, n=1.33, L_abs=None, keyed=False):
    """Return a sensor sphere (radius 0.3 m, index n) in a water world (an orb of 1.2 m, index n), sensor 0.

    ``rule`` "born": the sphere is a sensitive volume counting by volume (every entry a COUNTING row with ``eff``, and
    a photon emitted inside counted at its emission with that detection: ``born_row``); "entry": the same volume
    counting only entries (a photon born inside leaves uncounted); "exit": the sphere's border out into the water is a
    DIELECTRIC_RT row (R, T, eff) that detects, credited to the sphere (the volume left: ``credit_left``); "exit_off":
    that border without the crediting rule (the volume entered, water, holds no sensor).  ``keyed``: eff as the key
    surfaces/sensor/EFFICIENCY (shared by both rows)."""  # noqa: D205, D209
    kinds = [solids.orb(1.2, name="world"), solids.orb(0.3, name="sensor")]
    vols = [navarrays.Volume("world", 0, np.eye(3), np.zeros(3), -1, 0),
            navarrays.Volume("world/sensor", 1, np.eye(3), np.zeros(3), 0, 1)]
    nav = navarrays.build(vols, kinds, outside_cls=2)
    values = {"materials/Water/RINDEX": Table(E_KNOTS, np.array([n, n])),
              "materials/SensLike/RINDEX": Table(E_KNOTS, np.array([n, n]))}
    absorption = None
    if L_abs is not None:
        values["materials/Water/ABSLENGTH"] = Table(E_KNOTS, np.array([L_abs, L_abs]))
        absorption = "materials/Water/ABSLENGTH"
    media = [Medium("Water", "materials/Water/RINDEX", absorption), Medium("SensLike", "materials/SensLike/RINDEX")]
    if keyed:
        values["surfaces/sensor/EFFICIENCY"] = np.float64(eff)
        eff = "surfaces/sensor/EFFICIENCY"
    rows = [SurfaceRow("absorber", "ABSORBER"), SurfaceRow("dielectric", "DIELECTRIC"),
            SurfaceRow("entry", "COUNTING", {"eff": eff}),
            SurfaceRow("exit", "DIELECTRIC_RT", {"R": R, "T": T, "eff": eff})]
    row_of = np.full((3, 3), 0, np.int32)
    row_of[0, 1] = 2 if rule in ("born", "entry") else 1
    row_of[1, 0] = 3 if rule.startswith("exit") else 1
    credit_left = born_row = None
    if rule == "born":
        born_row = np.array([-1, 2, -1], np.int32)
    if rule == "exit":
        credit_left = np.zeros((3, 3), bool)
        credit_left[1, 0] = True
    return Detector(f"sensor_volume_{rule}", nav, media, np.array([0, 1, -1], np.int32), ["Water", "SensLike", "OUT"],
                    rows, row_of, np.array([-1, 0, -1], np.int32), ["sphere"], [SensorGroup("s", np.arange(1))],
                    Params(values), credit_left=credit_left, born_row=born_row)


def glisur_toy(kind="dielectric", n1=1.0, n2=1.5, polish=0.3, R=1.0, T=0.0, eff=0.0, keyed=False):
    """Return the GLISUR ring counter: a flat interface z = 0 between an upper medium (index n1, a world orb of 1.2
    m) and a lower box (|x|, |y| < 1.1 m, -1.1 < z < 0; index n2, or opaque for ``kind`` "metal") with a GLISUR row
    (GLISUR_DIELECTRIC or GLISUR_METAL) on the crossing down (and up, for the dielectric), inside a counting shell r
    in [1, 1.01] m cut into 36 polar rings of 5 degrees about +z by z slabs (rings 0-17 in the upper medium, 18-35
    in the box), each its own sensor absorbing and counting every photon that enters it.  A beam aimed at the origin
    counts the polar angle of its photons' last direction off the surface, the rings of Geant4's boundary runs.
    ``keyed``: polish (and R) as keys, which draws the facets from the cover mixture."""  # noqa: D205, D209
    kinds = [solids.orb(1.2, name="world"), solids.box(1.1, 1.1, 0.55, name="lower")]
    for i in range(36):  # the shell cut by the slab cos(theta_1) <= z <= cos(theta_0): planes, no cone apex at the
        # origin where every photon starts (a photon from the origin enters the shell at r = 1, z = cos(theta))
        z1, z0 = np.cos(np.deg2rad(5.0 * (i + 1))), np.cos(np.deg2rad(5.0 * i))
        shell = solids.boolean("sub", solids.orb(1.01), solids.orb(1.0), name=f"ring{i}")
        slab = solids.box(1.1, 1.1, 0.5 * (z0 - z1))
        kinds.append(solids.boolean("inter", shell, slab, t=np.array([0.0, 0.0, 0.5 * (z0 + z1)]), name=f"ring{i}"))
    vols = [navarrays.Volume("world", 0, np.eye(3), np.zeros(3), -1, 0),
            navarrays.Volume("world/lower", 1, np.eye(3), np.array([0.0, 0.0, -0.55]), 0, 2)]
    vols += [navarrays.Volume(f"world/ring{i}" if i < 18 else f"world/lower/ring{i}", 2 + i, np.eye(3), np.zeros(3),
                              0 if i < 18 else 1, 1) for i in range(36)]
    nav = navarrays.build(vols, kinds, outside_cls=3)
    values = {"materials/Upper/RINDEX": Table(E_KNOTS, np.array([n1, n1])),
              "materials/Lower/RINDEX": Table(E_KNOTS, np.array([n2, n2])),
              "materials/SensLike/RINDEX": Table(E_KNOTS, np.array([1.0, 1.0]))}
    media = [Medium("Upper", "materials/Upper/RINDEX"), Medium("Lower", "materials/Lower/RINDEX"),
             Medium("SensLike", "materials/SensLike/RINDEX")]
    cols = {"R": R, "polish": polish, "eff": eff}
    if keyed:
        values["surfaces/Glisur/REFLECTIVITY"] = np.float64(R)
        values["surfaces/Glisur/polish"] = np.float64(polish)
        cols.update(R="surfaces/Glisur/REFLECTIVITY", polish="surfaces/Glisur/polish")
    if kind == "dielectric":
        cols["T"] = T
        row = SurfaceRow("Glisur", "GLISUR_DIELECTRIC", cols)
    else:
        row = SurfaceRow("Glisur", "GLISUR_METAL", cols)
    rows = [SurfaceRow("ring", "SPECULAR_METAL", {"R": 0.0, "eff": 1.0}), SurfaceRow("absorber", "ABSORBER"), row]
    row_of = np.full((4, 4), 1, np.int32)
    row_of[0, 1] = row_of[2, 1] = 0
    row_of[0, 2] = 2
    if kind == "dielectric":
        row_of[2, 0] = 2
    class_medium = np.array([0, 2, 1 if kind == "dielectric" else -1, -1], np.int32)
    sensor_of_vol = np.array([-1, -1] + list(range(36)) + [-1], np.int32)
    return Detector(f"glisur_{kind}", nav, media, class_medium, ["Upper", "SensLike", "Lower", "OUT"], rows, row_of,
                    sensor_of_vol, [f"ring{i}" for i in range(36)], [SensorGroup("rings", np.arange(36))],
                    Params(values))


def glisur_beam(theta_deg, lam=420.0):
    """A beam from 0.5 m up the incidence line onto the origin at ``theta_deg`` off the normal +z (unpolarised: a
    random linear state per photon, Geant4's random polarisation)."""  # noqa: D205, D209
    from lumen_amd import sources

    t = np.deg2rad(theta_deg)
    return sources.Point((-0.5 * np.sin(t), 0.0, 0.5 * np.cos(t)), lam=lam, n_photons=1.0,
                         direction=(np.sin(t), 0.0, -np.cos(t)))


MIE_CASES = {"A": (0.6, 0.3, 0.7), "B": (0.0, 0.8, 0.5), "water": (0.95, 0.3, 1.0)}  # (g_f, g_b, r)
MIE_R, MIE_L, MIE_N, MIE_E = 1.0, 20.0, 1.33, 3.0  # sphere radius (m), MIEHG (m), RINDEX, photon energy (eV)


def mie_toy(case="A", rings=1, length=MIE_L):
    """Return the Mie toy of the reference's Mie transport tests: a Mie-only medium (MIEHG ``length``, RINDEX 1.33,
    no absorption) in a world orb of 1.2 m, inside a counting shell r in [1, 1.01] m cut into ``rings`` polar rings
    of equal angle (each its own sensor, every photon absorbed and counted where it enters one), the lobe constants
    of ``case`` (MIE_CASES) as parameter keys."""  # noqa: D205, D209
    gf, gb, r = MIE_CASES[case]
    d = np.pi / rings
    kinds = [solids.orb(1.2, name="world")] + [solids.sphere(MIE_R, 1.01, i * d, d, name=f"ring{i}")
                                              for i in range(rings)]
    vols = [navarrays.Volume("world", 0, np.eye(3), np.zeros(3), -1, 0)]
    vols += [navarrays.Volume(f"world/ring{i}", 1 + i, np.eye(3), np.zeros(3), 0, 1) for i in range(rings)]
    nav = navarrays.build(vols, kinds, outside_cls=2)
    m = "materials/MieLike/"
    values = {m + "RINDEX": Table(E_KNOTS, np.array([MIE_N, MIE_N])),
              m + "MIEHG": Table(E_KNOTS, np.array([length, length])),
              m + "MIEHG_FORWARD": np.float64(gf), m + "MIEHG_BACKWARD": np.float64(gb),
              m + "MIEHG_FORWARD_RATIO": np.float64(r),
              "materials/SensLike/RINDEX": Table(E_KNOTS, np.array([MIE_N, MIE_N]))}
    media = [Medium("MieLike", m + "RINDEX", mie_length=m + "MIEHG", mie_forward=m + "MIEHG_FORWARD",
                    mie_backward=m + "MIEHG_BACKWARD", mie_ratio=m + "MIEHG_FORWARD_RATIO"),
             Medium("SensLike", "materials/SensLike/RINDEX")]
    rows = [SurfaceRow("ring", "SPECULAR_METAL", {"R": 0.0, "eff": 1.0}), SurfaceRow("absorber", "ABSORBER")]
    row_of = np.full((3, 3), -1, np.int32)
    row_of[0, 1] = 0
    row_of[0, 2] = row_of[1, 0] = row_of[1, 2] = 1
    sensor_of_vol = np.array([-1] + list(range(rings)) + [-1], np.int32)
    return Detector(f"mie_{case}_{rings}", nav, media, np.array([0, 1, -1], np.int32), ["MieLike", "SensLike", "OUT"],
                    rows, row_of, sensor_of_vol, [f"ring{i}" for i in range(rings)],
                    [SensorGroup("rings", np.arange(rings))], Params(values))


def cone_toy(rings=90, n=MIE_N):
    """Return the Cherenkov cone toy: plain water (RINDEX n flat over 1.5-5 eV, nothing absorbs or scatters) in a
    world orb of 1.2 m inside a counting shell r in [1, 1.01] m cut into ``rings`` polar rings of equal angle, each
    its own sensor absorbing and counting every photon that enters it."""  # noqa: D205, D209
    d = np.pi / rings
    kinds = [solids.orb(1.2, name="world")] + [solids.sphere(MIE_R, 1.01, i * d, d, name=f"ring{i}")
                                              for i in range(rings)]
    vols = [navarrays.Volume("world", 0, np.eye(3), np.zeros(3), -1, 0)]
    vols += [navarrays.Volume(f"world/ring{i}", 1 + i, np.eye(3), np.zeros(3), 0, 1) for i in range(rings)]
    nav = navarrays.build(vols, kinds, outside_cls=2)
    values = {"materials/Water/RINDEX": Table(E_KNOTS, np.array([n, n])),
              "materials/SensLike/RINDEX": Table(E_KNOTS, np.array([n, n]))}
    media = [Medium("Water", "materials/Water/RINDEX"), Medium("SensLike", "materials/SensLike/RINDEX")]
    rows = [SurfaceRow("ring", "SPECULAR_METAL", {"R": 0.0, "eff": 1.0}), SurfaceRow("absorber", "ABSORBER")]
    row_of = np.full((3, 3), -1, np.int32)
    row_of[0, 1] = 0
    row_of[0, 2] = row_of[1, 0] = row_of[1, 2] = 1
    sensor_of_vol = np.array([-1] + list(range(rings)) + [-1], np.int32)
    return Detector(f"cone_{rings}", nav, media, np.array([0, 1, -1], np.int32), ["Water", "SensLike", "OUT"],
                    rows, row_of, sensor_of_vol, [f"ring{i}" for i in range(rings)],
                    [SensorGroup("rings", np.arange(rings))], Params(values))


def cone_rings_closed_form(rings, z0, z1, cos_c, beta, n, R=MIE_R, m=200001):
    """Return (share, mean time ns) per ring of the cone toy for a track along +z from z0 to z1 (m) started at t = 0:
    a photon born at z (uniform on the track, at (z - z0) / (beta c)) along the cone's polar angle reaches the sphere of
    radius R at the polar angle alpha(z) of z + t cos, t the root of t^2 + 2 z cos t + z^2 - R^2 = 0, after the
    time t n / c (the group velocity c / n of a flat index); the azimuth plays no part.  Midpoint rule on m points."""  # noqa: D205, D209
    z = z0 + (np.arange(m) + 0.5) / m * (z1 - z0)
    t = -z * cos_c + np.sqrt(z * z * cos_c * cos_c - z * z + R * R)
    alpha = np.arccos(np.clip((z + t * cos_c) / R, -1.0, 1.0))
    ring = np.minimum((alpha / (np.pi / rings)).astype(np.int64), rings - 1)
    arrive = (z - z0) / (beta * units.C_VACUUM) + t * n / units.C_VACUUM
    share = np.bincount(ring, minlength=rings) / m
    tsum = np.bincount(ring, weights=arrive, minlength=rings) / m
    return share, np.where(share > 0, tsum / np.where(share > 0, share, 1.0), 0.0)


def mie_hg_cdf_angle(theta, g):
    """P(scattering angle <= theta) of HG(g), float64 (the textbook CDF in 1 - cos theta = 2 sin^2(theta/2))."""
    s2 = np.sin(0.5 * np.asarray(theta, np.float64)) ** 2
    if g == 0.0:
        return s2
    q = (1.0 - g) ** 2 + 4.0 * g * s2
    return (1.0 - g * g) / (2.0 * g) * (1.0 / (1.0 - g) - 1.0 / np.sqrt(q))


def mie_cdf_angle(theta, gf, gb, r):
    """P(scattering angle <= theta) of the mixture r HG(g_f) + (1 - r) HG(-g_b)."""
    return r * mie_hg_cdf_angle(theta, gf) + (1.0 - r) * mie_hg_cdf_angle(theta, -gb)


def specular_closed_form(a, R, L):
    """Q/N of the specular sensor sphere: (1 - R) e^{-a/L} / (1 - R e^{-2a/L})."""
    return (1 - R) * np.exp(-a / L) / (1 - R * np.exp(-2 * a / L))


def lambert_closed_form(a, R, L):
    """Q/N of the Lambertian sensor sphere: (1 - R) e^{-a/L} / (1 - R 2 (1 - (1 + b) e^{-b}) / b^2), b = 2a/L."""
    b = 2 * a / L
    return (1 - R) * np.exp(-a / L) / (1 - R * 2 * (1 - (1 + b) * np.exp(-b)) / b**2)


def group_velocity(n):
    """The group velocity of a flat index table (m/ns): c / n."""
    return units.C_VACUUM / n


# a GDML of the added solids (cone, tube with a phi cut, polycone, ellipsoid, trd) in a Mie liquid: the reader's test
GDML_SOLIDS = """<?xml version="1.0" encoding="UTF-8"?>
<gdml>
<define>
  <matrix name="RIDX" coldim="2" values="1.0*eV 1.33 12.0*eV 1.33"/>
  <matrix name="MIEL" coldim="2" values="1.0*eV 20*m 12.0*eV 20*m"/>
  <matrix name="MIEF" coldim="1" values="0.6"/>
  <matrix name="MIEB" coldim="1" values="0.3"/>
  <matrix name="MIER" coldim="1" values="0.7"/>
</define>
<materials>
  <material name="Liq"><property name="RINDEX" ref="RIDX"/><property name="MIEHG" ref="MIEL"/>
    <property name="MIEHG_FORWARD" ref="MIEF"/><property name="MIEHG_BACKWARD" ref="MIEB"/>
    <property name="MIEHG_FORWARD_RATIO" ref="MIER"/></material>
</materials>
<solids>
  <box name="WorldBox" x="4" y="4" z="4" lunit="m"/>
  <cone name="C" rmin1="100" rmax1="500" rmin2="150" rmax2="300" z="800" startphi="0" deltaphi="360" aunit="deg" lunit="mm"/>
  <tube name="T" rmin="0" rmax="300" z="400" startphi="30" deltaphi="120" aunit="deg" lunit="mm"/>
  <polycone name="P" startphi="0" deltaphi="360" aunit="deg" lunit="mm">
    <zplane z="-300" rmin="0" rmax="200"/><zplane z="0" rmin="50" rmax="300"/><zplane z="200" rmin="0" rmax="100"/>
  </polycone>
  <ellipsoid name="E" ax="200" by="300" cz="400" zcut1="-300" zcut2="350" lunit="mm"/>
  <trd name="D" x1="400" x2="200" y1="300" y2="500" z="600" lunit="mm"/>
</solids>
<structure>
  <volume name="cV"><materialref ref="Liq"/><solidref ref="C"/></volume>
  <volume name="tV"><materialref ref="Liq"/><solidref ref="T"/></volume>
  <volume name="pV"><materialref ref="Liq"/><solidref ref="P"/></volume>
  <volume name="eV"><materialref ref="Liq"/><solidref ref="E"/></volume>
  <volume name="dV"><materialref ref="Liq"/><solidref ref="D"/></volume>
  <volume name="World"><materialref ref="Liq"/><solidref ref="WorldBox"/>
    <physvol name="pc"><volumeref ref="cV"/><position name="p1" x="-1" y="-1" z="0" unit="m"/></physvol>
    <physvol name="pt"><volumeref ref="tV"/><position name="p2" x="1" y="-1" z="0" unit="m"/></physvol>
    <physvol name="pp"><volumeref ref="pV"/><position name="p3" x="-1" y="1" z="0" unit="m"/></physvol>
    <physvol name="pe"><volumeref ref="eV"/><position name="p4" x="1" y="1" z="0" unit="m"/></physvol>
    <physvol name="pd"><volumeref ref="dV"/><position name="p5" x="0" y="0" z="1.2" unit="m"/></physvol>
  </volume>
</structure>
<setup name="Default" version="1.0"><world ref="World"/></setup>
</gdml>
"""


def gdml_framed() -> str:
    """A world of liquid holding 16 thin plates sharing one rotation (45 degrees about z: their world boxes are 6x
    their own, a frame group), 3 plates tilted 30 degrees about x (too few for a group) and 4 orb sensors."""
    pv = []
    for i, (x, y) in enumerate((x, y) for x in (-1.2, -0.4, 0.4, 1.2) for y in (-1.2, -0.4, 0.4, 1.2)):
        pv.append(f'<physvol name="plate{i}"><volumeref ref="plV"/><position name="pp{i}" x="{x}" y="{y}" z="0" '
                  f'unit="m"/><rotationref ref="r45"/></physvol>')
    for i, x in enumerate((-1.0, 0.0, 1.0)):
        pv.append(f'<physvol name="tilt{i}"><volumeref ref="plV"/><position name="pt{i}" x="{x}" y="0.3" z="1.2" '
                  f'unit="m"/><rotationref ref="rx30"/></physvol>')
    for i, (x, y) in enumerate(((-0.8, -0.8), (0.8, -0.8), (-0.8, 0.8), (0.8, 0.8))):
        pv.append(f'<physvol name="ball{i}"><volumeref ref="oV"/><position name="po{i}" x="{x}" y="{y}" z="-1.2" '
                  f'unit="m"/></physvol>')
    body = "\n    ".join(pv)
    return f"""<?xml version="1.0" encoding="UTF-8"?>
<gdml>
<define>
  <matrix name="RIDX" coldim="2" values="1.0*eV 1.33 12.0*eV 1.33"/>
  <matrix name="RIDX2" coldim="2" values="1.0*eV 1.50 12.0*eV 1.50"/>
  <matrix name="ABSL" coldim="2" values="1.0*eV 5*m 12.0*eV 5*m"/>
  <matrix name="CREF" coldim="2" values="1.0*eV 0.0 12.0*eV 0.0"/>
  <matrix name="CEFF" coldim="2" values="1.0*eV 1.0 12.0*eV 1.0"/>
  <rotation name="r45" z="45" unit="deg"/>
  <rotation name="rx30" x="30" unit="deg"/>
</define>
<materials>
  <material name="Liq"><property name="RINDEX" ref="RIDX"/><property name="ABSLENGTH" ref="ABSL"/></material>
  <material name="Glass"><property name="RINDEX" ref="RIDX2"/><property name="ABSLENGTH" ref="ABSL"/></material>
</materials>
<solids>
  <box name="WorldBox" x="4" y="4" z="4" lunit="m"/>
  <box name="Plate" x="400" y="20" z="20" lunit="mm"/>
  <orb name="O" r="100" lunit="mm"/>
  <opticalsurface name="Cathode" model="unified" finish="polished" type="dielectric_metal" value="0">
    <property name="REFLECTIVITY" ref="CREF"/><property name="EFFICIENCY" ref="CEFF"/>
  </opticalsurface>
</solids>
<structure>
  <volume name="plV"><materialref ref="Glass"/><solidref ref="Plate"/></volume>
  <volume name="oV"><materialref ref="Glass"/><solidref ref="O"/></volume>
  <volume name="World"><materialref ref="Liq"/><solidref ref="WorldBox"/>
    {body}
  </volume>
  <skinsurface name="cathodes" surfaceproperty="Cathode"><volumeref ref="oV"/></skinsurface>
</structure>
<setup name="Default" version="1.0"><world ref="World"/></setup>
</gdml>
"""


GDML_TORI = """<?xml version="1.0" encoding="UTF-8"?>
<gdml>
<define>
  <matrix name="RIDX" coldim="2" values="1.0*eV 1.33 12.0*eV 1.33"/>
  <matrix name="RIDX2" coldim="2" values="1.0*eV 1.50 12.0*eV 1.50"/>
  <matrix name="ABSL" coldim="2" values="1.0*eV 3*m 12.0*eV 3*m"/>
  <rotation name="rx" x="60" y="20" unit="deg"/>
</define>
<materials>
  <material name="Liq"><property name="RINDEX" ref="RIDX"/></material>
  <material name="Glass"><property name="RINDEX" ref="RIDX2"/><property name="ABSLENGTH" ref="ABSL"/></material>
</materials>
<solids>
  <box name="WorldBox" x="4" y="4" z="4" lunit="m"/>
  <torus name="T1" rmin="0" rmax="200" rtor="500" startphi="0" deltaphi="360" aunit="deg" lunit="mm"/>
  <torus name="T2" rmin="60" rmax="150" rtor="400" startphi="30" deltaphi="240" aunit="deg" lunit="mm"/>
  <torus name="T3" rmin="0" rmax="300" rtor="400" startphi="0" deltaphi="360" aunit="deg" lunit="mm"/>
</solids>
<structure>
  <volume name="t1V"><materialref ref="Liq"/><solidref ref="T1"/></volume>
  <volume name="t2V"><materialref ref="Glass"/><solidref ref="T2"/></volume>
  <volume name="t3V"><materialref ref="Glass"/><solidref ref="T3"/></volume>
  <volume name="World"><materialref ref="Liq"/><solidref ref="WorldBox"/>
    <physvol name="p1"><volumeref ref="t1V"/><position name="q1" x="-0.9" y="-0.9" z="0" unit="m"/></physvol>
    <physvol name="p2"><volumeref ref="t2V"/><position name="q2" x="0.9" y="0.9" z="0" unit="m"/>
      <rotationref ref="rx"/></physvol>
    <physvol name="p3"><volumeref ref="t3V"/><position name="q3" x="0.9" y="-0.9" z="0.8" unit="m"/></physvol>
  </volume>
</structure>
<setup name="Default" version="1.0"><world ref="World"/></setup>
</gdml>
"""
