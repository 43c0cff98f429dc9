"""Pin the oracle (the FP64 CPU restatement) against what fixes it independently of its own formulas: published
known-answer vectors, Geant4's histograms, closed forms, textbook optics and the reference tutorial's numbers."""

import json
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest
from scipy import stats

import lumen_amd as L
from lumen_amd import pack, units
from lumen_amd.detector import example_tank, water_tank
from oracle import oracle as O
from toys import lambert_closed_form, specular_closed_form, sphere_toy

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P_MIN = 1e-3  # one statistical assertion's significance


# --------------------------------------------------------------------------------------------- random numbers
KAT = [  # Random123's published known-answer vectors for philox4x32_10 (kat_vectors): counter, key, output
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert tuple(int(x) for x in O.philox(ctr, key)) == out


def test_philox_three_implementations_agree():
    """The oracle's Philox, the engine's header (compiled for the host) and rocRAND's philox4x32_10 engine agree on
    random counters and keys (rocRAND is an independent implementation of the same published generator)."""
    src = r"""
#include <cstdio>
#include <cstdlib>
#include <rocrand/rocrand_philox4x32_10.h>
#include "philox.h"
struct X : rocrand_device::philox4x32_10_engine { uint4 tr(uint4 c, uint2 k) { return ten_rounds(c, k); } };
int main() {
  X x; unsigned s = 12345u; int bad = 0;
  for (int i = 0; i < 2000; ++i) {
    unsigned v[6];
    for (int j = 0; j < 6; ++j) { s = s * 1664525u + 1013904223u; v[j] = s ^ (s >> 7); }
    uint4 r = x.tr(uint4{v[0], v[1], v[2], v[3]}, uint2{v[4], v[5]});
    lh_u4 e = lh_philox(lh_u4{v[0], v[1], v[2], v[3]}, v[4], v[5]);
    if (r.x != e.x || r.y != e.y || r.z != e.z || r.w != e.w) ++bad;
    printf("%u %u %u %u %u %u %u %u %u %u\n", v[0], v[1], v[2], v[3], v[4], v[5], e.x, e.y, e.z, e.w);
  }
  fprintf(stderr, "mismatch %d\n", bad);
  return bad != 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "kat.hip")
        open(f, "w").write(src)
        exe = os.path.join(d, "kat")
        r = subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-I", os.path.join(ROOT, "lumen_amd", "csrc"), f, "-o", exe],
                           capture_output=True, text=True)
        if r.returncode != 0:
            pytest.skip(f"hipcc could not build the host check: {r.stderr[-300:]}")
        run = subprocess.run([exe], capture_output=True, text=True)
        assert run.returncode == 0, run.stderr
        rows = np.array([[int(x) for x in ln.split()] for ln in run.stdout.strip().splitlines()], np.uint64)
    for row in rows[:200]:
        assert tuple(int(x) for x in O.philox(row[:4], row[4:6])) == tuple(int(x) for x in row[6:])


# --------------------------------------------------------------------------------------------- media
def test_water_rayleigh_closed_form():
    """Geant4's Water Rayleigh rule on the example tank's RINDEX knots (Einstein-Smoluchowski, computed in the
    survey of the reference from rules/optics.py:189-226): 3184, 399.3, 96.6, 19.65 m; 217.8 m at 400 nm."""
    det = example_tank()
    rt = det.params()["materials/Water/RINDEX"]
    Lr = pack.water_rayleigh(rt.energy, rt.value)
    np.testing.assert_allclose(Lr, [3184.0, 399.3, 96.6, 19.65], rtol=2e-3)
    at400 = np.interp(units.HC / 400.0, rt.energy, Lr)
    assert abs(at400 - 217.8) < 0.2


def test_water_at_400nm():
    """The example tank's water at 400 nm: n = 1.3414, ABSLENGTH 54.02 m (tables linear in photon energy)."""
    det = example_tank()
    p = det.params()
    E = units.HC / 400.0
    n = np.interp(E, p["materials/Water/RINDEX"].energy, p["materials/Water/RINDEX"].value)
    La = np.interp(E, p["materials/Water/ABSLENGTH"].energy, p["materials/Water/ABSLENGTH"].value)
    assert abs(n - 1.3414) < 1e-4 and abs(La - 54.02) < 0.01


def test_group_index_converges_to_the_group_index():
    """Geant4's GROUPVEL rule on dense knots of a smooth n(E) converges to the group index n + E dn/dE."""
    e = np.linspace(2.0, 4.2, 400)
    n_of = lambda x: 1.30 + 0.01 * x + 0.004 * x**2  # noqa: E731
    eg, ng = pack.group_index_knots(e, n_of(e))
    x = np.array([2.5, 3.1, 3.9])
    exact = n_of(x) + x * (0.01 + 0.008 * x)
    got = 1.0 / np.interp(x, eg, 1.0 / ng)
    np.testing.assert_allclose(got, exact, rtol=2e-5)


# --------------------------------------------------------------------------------------------- navigator referee
def _ray_sphere(o, d, c, r):
    oc = o - c
    b = oc @ d
    disc = b * b - (oc @ oc - r * r)
    if disc < 0:
        return math.inf, math.inf
    s = math.sqrt(disc)
    return -b - s, -b + s


def test_navigator_referee_closed_forms_example_tank():
    """Rays to the top PMT's face (a plane, 0.37 m from the source point), a side PMT's face from the centre, the
    tank's wall along a diagonal and past a PMT."""
    det = example_tank()
    o = O.OracleRun(det)
    x = np.array([[0, 0, 0.1], [0, 0, 0.0], [0, 0, 0.1], [0.1, 0.2, 0.3]])
    d = np.array([[0, 0, 1.0], [1.0, 0, 0], [1.0, 1.0, 0] / np.sqrt(2), [0, -1.0, 0]])
    t, va, nrm, _ = o.query(x, d, 0)
    # the arrays hold the geometry in float32 (the navigator's precision): 1e-7 is its rounding, not the referee's
    np.testing.assert_allclose(t, [0.37, 0.47, (0.5) * np.sqrt(2), 0.7], rtol=1e-7)
    assert list(va) == [2, 4, 9, 9]  # PMT_1 (volume 2), PMT_3 (volume 4), the outside (9) twice
    np.testing.assert_allclose(nrm[0], [0, 0, -1], atol=1e-12)


def test_navigator_referee_closed_forms_water_tank():
    """Rays to the barrel (a cylinder), to sphere-cap domes (spheres) and to a cap's base plane."""
    det = water_tank()
    o = O.OracleRun(det)
    rng = np.random.default_rng(0)
    R = det.nav.host["R"].astype(np.float64)
    T = det.nav.host["t"].astype(np.float64)
    n_ok = 0
    for k in rng.choice(np.arange(1, det.nav.n), 64, replace=False):
        axis = R[k][:, 2]
        c = T[k]  # the cap's sphere centre
        top = c + 0.10 * axis  # the dome's apex
        start = top + 0.5 * axis + 0.02 * np.cross(axis, [0.3, 0.4, 0.5])
        d = (top - start) / np.linalg.norm(top - start)
        t, va, _, _ = o.query(start[None], d[None], 0)
        t0, _ = _ray_sphere(start, d, c, 0.10)
        if va[0] == k:
            assert abs(t[0] - t0) < 1e-7 * max(1.0, t0)  # float32-stored radii and poses
            n_ok += 1
    assert n_ok >= 60
    # along +z from the axis at z = 0: the top endcap's central PMT dome, else the top lid
    t, va, _, _ = o.query(np.array([[0.0, 0.0, 0.0]]), np.array([[0.0, 0.0, 1.0]]), 0)
    assert va[0] > 0 and abs(t[0] - (1.5 - 0.003 - 0.04)) < 1e-7


# --------------------------------------------------------------------------------------------- surfaces and lobes
def _chi2_p(a, b):
    a, b = np.ravel(a).astype(float), np.ravel(b).astype(float)
    n1, n2 = a.sum(), b.sum()
    use = (a + b) > 0
    c2 = float(np.sum((a[use] / n1 - b[use] / n2) ** 2 / ((a[use] + b[use]) / (n1 * n2))))
    return float(stats.chi2.sf(c2, int(use.sum()) - 1))


MANIFEST = json.load(open(os.path.join(GOLDEN, "g4_unified_manifest.json")))["cases"]


@pytest.mark.parametrize("name", sorted(MANIFEST))
def test_facet_path_is_geant4s(name):
    """The unified facet path against Geant4 11.3.0 (R = 1, lobe constant 1): the last direction's (dz, |phi|)
    histogram, two-sample chi2 (the reference's tests/unit/test_g4_boundary.py:61-77)."""
    case = MANIFEST[name]
    th = math.radians(case["theta"])
    wi = np.array([math.sin(th), 0.0, -math.cos(th)])
    wo = O.facet_paths(case["value"], wi, 1 << 20, seed=case["seed"])
    with np.load(os.path.join(GOLDEN, case["file"])) as z:
        h = np.histogram2d(wo[:, 2], np.abs(np.arctan2(wo[:, 1], wo[:, 0])), bins=[z["edges_z"], z["edges_phi"]])[0]
        assert _chi2_p(z["hist"], h) > P_MIN


def test_rayleigh_and_lambert_densities():
    """Rayleigh from a linear state er: c = k . er has density 3/4 (1 - c^2); Lambert about +z: cos^2 uniform."""
    ray, lam = O.lobes([1.0, 0.0, 0.0], [0.0, 0.0, 0.0], 400000, seed=3)
    c = ray[:, 0]
    edges = np.linspace(-1, 1, 21)
    obs = np.histogram(c, edges)[0]
    cdf = lambda x: 0.75 * (x - x**3 / 3) + 0.5  # noqa: E731
    exp = np.diff(cdf(edges)) * len(c)
    assert stats.chisquare(obs, exp).pvalue > P_MIN
    np.testing.assert_allclose(np.linalg.norm(ray, axis=1), 1.0, atol=1e-12)
    obs = np.histogram(lam[:, 2] ** 2, np.linspace(0, 1, 21))[0]
    assert stats.chisquare(obs, np.full(20, len(lam) / 20)).pvalue > P_MIN
    assert np.all(lam[:, 2] > 0)


def test_fresnel_textbook():
    n1, n2 = 1.0, 1.5
    rs, rp, ts, tp = O.fresnel(1.0, n1, n2)
    assert abs(rs - (n1 - n2) / (n1 + n2)) < 1e-12 and abs(rp + rs) < 1e-12
    cb = math.cos(math.atan(n2 / n1))  # Brewster
    assert abs(O.fresnel(cb, n1, n2)[1]) < 1e-12
    for c in np.linspace(0.05, 1.0, 12):  # energy: R + T = 1 for each of s and p
        rs, rp, ts, tp = O.fresnel(c, n1, n2)
        q = math.sqrt(n2 * n2 - n1 * n1 * (1 - c * c))
        assert abs(abs(rs) ** 2 + q / (n1 * c) * abs(ts) ** 2 - 1) < 1e-12
        assert abs(abs(rp) ** 2 + q / (n1 * c) * abs(tp) ** 2 - 1) < 1e-12
    for c in (0.1, 0.3, 0.5):  # total internal reflection beyond the critical angle (1.5 -> 1.0, c < 0.745)
        rs, rp, _, _ = O.fresnel(c, 1.5, 1.0)
        assert abs(abs(rs) - 1) < 1e-12 and abs(abs(rp) - 1) < 1e-12


# --------------------------------------------------------------------------------------------- sources
def test_cherenkov_yield_frank_tamm():
    """With a flat index every photon's weight is Frank-Tamm's yield over samples: N/L = 2 pi alpha z^2
    (1/lam_min - 1/lam_max) (1 - 1/(beta n)^2)."""
    det = sphere_toy(a=2.0, R=0.0, L_abs=None)
    o = O.OracleRun(det)
    src = L.sources.CherenkovSegment("Water", start=(0, 0, -0.5), direction=(0, 0, 1), length=1.0, beta=1.0)
    _, h = o.trace(src, 4096, seed=1)
    lam_min, lam_max = units.HC / 5.0, units.HC / 1.5
    n = 1.33
    N = 2 * math.pi * units.FINE_STRUCTURE * (1 / lam_min - 1 / lam_max) * 1e9 * (1 - 1 / n**2)
    assert abs(h[14] - N) / N < 1e-5


# --------------------------------------------------------------------------------------------- the whole loop
def test_specular_sphere_closed_form():
    """Every photon's light in the specular sensor sphere is the closed form (1-R)e^{-a/L}/(1-R e^{-2a/L})."""
    a, R, La = 1.0, 0.5, 10.0
    det = sphere_toy(a=a, R=R, L_abs=La)
    o = O.OracleRun(det)
    tal, h = o.trace(L.sources.Point((0, 0, 0), lam=400.0, n_photons=1.0), 4096, seed=2,
                     controls=L.Controls(roulette=1e-12))
    assert abs(tal[0, 0] - specular_closed_form(a, R, La)) < 3e-5 * specular_closed_form(a, R, La)
    assert abs(h[14] - h[:9].sum()) < 1e-9  # the health balance


def test_lambert_sphere_closed_form():
    a, R, La = 1.0, 0.6, 4.0
    det = sphere_toy(a=a, surface="PAINTED", R=R, L_abs=La)
    o = O.OracleRun(det)
    n = 1 << 16
    tal, _ = o.trace(L.sources.Point((0, 0, 0), lam=400.0, n_photons=1.0), n, seed=3)
    s, q2 = tal[0, 0], tal[0, 2]
    err = math.sqrt(max(n / (n - 1) * (q2 - s * s / n), 0))
    assert abs(s - lambert_closed_form(a, R, La)) < 4 * err


def test_enclosing_sensor_with_rayleigh_counts_everything():
    """A sensor enclosing a scattering, non-absorbing medium counts every photon: Q/N = 1 (test_toys.py:162-172)."""
    det = sphere_toy(a=1.0, R=0.0, L_abs=None, L_ray=0.3)
    o = O.OracleRun(det)
    n = 1 << 15
    tal, h = o.trace(L.sources.Point((0, 0, 0), lam=400.0, n_photons=1.0), n, seed=4)
    s, q2 = tal[0, 0], tal[0, 2]
    err = math.sqrt(max(n / (n - 1) * (q2 - s * s / n), 0))
    assert abs(s - 1.0) < max(4 * err, 1e-6)
    assert h[12] > 3 * n  # the photons did scatter


def test_group_velocity_timing():
    """With nothing reflected the mean arrival time is a / v_g (test_toys.py:175-182)."""
    det = sphere_toy(a=1.5, R=0.0, L_abs=None, n=1.40)
    o = O.OracleRun(det)
    tal, _ = o.trace(L.sources.Point((0, 0, 0), lam=400.0, n_photons=1.0), 1024, seed=5)
    assert abs(tal[0, 1] / tal[0, 0] / (1.5 / (units.C_VACUUM / 1.40)) - 1) < 1e-7  # the index is stored in float32


def test_tutorial_arrivals():
    """The example tank at the tutorial's settings against the published arrivals (docs/tutorial/first-hour.md):
    chi2 of the pulls (errors in quadrature), mean times and PE per arrival."""
    tut = json.load(open(os.path.join(GOLDEN, "tutorial_example_tank.json")))
    det = example_tank()
    o = O.OracleRun(det)
    n = 1 << 18
    tal, h = o.trace(L.sources.Point((0.0, 0.0, 0.1), lam=400.0, n_photons=2e5), n, seed=0)
    r = O.light_of(tal, h, 8, len(tal[0]) // 8 - 3, n)
    pub, perr = np.array(tut["arrivals"]), np.array(tut["arrivals_error"])
    pulls = (r["arrivals"] - pub) / np.hypot(r["arrivals_error"], perr)
    assert stats.chi2.sf(np.sum(pulls**2), 8) > 0.01
    np.testing.assert_allclose(r["mean_arrival_time"], tut["mean_arrival_time_ns"], atol=0.02)
    qe = det.params()["sensors/pmt/qe"]
    pe = (r["charge"] * qe.value[None, :]).sum(1)
    np.testing.assert_allclose(pe / r["arrivals"], 0.245, rtol=1e-6)
    np.testing.assert_allclose(r["arrivals_error"] / pub, perr / pub, rtol=0.15)  # the errors' size, not a pull


def test_moment_is_each_photons_square():
    """arrival2 is the sum over photons of the square of each photon's summed deposits: brute force, photon by photon
    (first_id makes photon i the same alone as in the run; tests/unit/test_transport.py:306-328)."""
    det = water_tank()
    o = O.OracleRun(det)
    x0 = (0.3, -0.2, 0.1)
    src = L.sources.Point(x0, lam=420.0, n_photons=100.0)
    n = 300
    tal, _ = o.trace(src, n, seed=9)
    S = det.n_sensors
    brute = np.zeros(S)
    total = np.zeros(S)
    for i in range(n):
        # photon i alone: one photon with the run's weight (n_photons / n) and the id i
        one, _ = o.trace(L.sources.Point(x0, lam=420.0, n_photons=100.0 / n), 1, seed=9, first_id=i)
        brute += one[0, :S] ** 2
        total += one[0, :S]
    np.testing.assert_allclose(total, tal[0, :S], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(brute, tal[0, 2 * S:3 * S], rtol=1e-12, atol=1e-15)


def test_bvh_walk_equals_brute_force():
    """The oracle's BVH walk (the reference's navigator, used for the work model and the CPU baseline) finds the
    same crossings as its brute-force referee: the same photons give the same tallies."""
    det = water_tank()
    o = O.OracleRun(det)
    x0, d = L.detector.muon_track()
    src = L.sources.CherenkovSegment("Water", start=tuple(x0), direction=tuple(d), length=1.0)
    a, ha = o.trace(src, 1 << 12, seed=3, bvh=True)
    b, hb = o.trace(src, 1 << 12, seed=3, bvh=False)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)
    assert ha[12] == hb[12]


def _framed_detector(tmp_path):
    from lumen_amd.detector import DATA
    from lumen_amd.detector.gdml import Group, Recipe, load
    from toys import gdml_framed

    f = tmp_path / "framed.gdml"
    f.write_text(gdml_framed())
    return load(str(f), Recipe("framed", "World", (Group("ball", "oV", values=str(DATA / "toy_pmt.json")),)))


def test_frame_group_walk_equals_brute_force(tmp_path):
    """Daughters sharing a rotation get a BVH in that rotation's frame (tighter boxes); the walk through it finds
    the brute-force referee's crossings: the same photons give the same tallies."""
    det = _framed_detector(tmp_path)
    h = det.nav.host
    assert det.nav.stats["framed"] == 1 and h["vg1"][0] - h["vg0"][0] == 2  # the plates' group, then the world's
    Rg = np.asarray(h["gR"][h["vg0"][0]], np.float64)
    assert abs(Rg[0, 0] - math.sqrt(0.5)) < 1e-6 and abs(Rg[2, 2] - 1.0) < 1e-6
    o = O.OracleRun(det)
    src = L.sources.Point((0.05, 0.1, 0.6), lam=420.0, n_photons=1.0)
    a, ha = o.trace(src, 1 << 12, seed=5, bvh=True)
    b, hb = o.trace(src, 1 << 12, seed=5, bvh=False)
    assert a[0, :4].sum() > 0
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)
    assert ha[12] == hb[12]


def _timed_sphere(t0, dt, nt, a=1.5, n_idx=1.40):
    return sphere_toy(a=a, R=0.0, L_abs=None, n=n_idx).timed(t0, dt, nt)


def test_time_histogram_reproduces_cubic_moments():
    """Every photon arrives at T = a / v_g: the cubic Lagrange split of a point on the nodes keeps the moments of
    order 0 to 3 (readout/grid.py:190-199), so the node masses sum to the arrivals and their first three time moments
    are T, T^2, T^3; nothing overflows."""
    a, n_idx = 1.5, 1.40
    T = a / (units.C_VACUUM / n_idx)
    det = _timed_sphere(0.0, 0.37, 40, a, n_idx)
    tal, h = O.OracleRun(det).trace(L.sources.Point((0, 0, 0), lam=400.0, n_photons=1.0), 512, seed=2)
    o = O.light_of(tal, h, 1, 1, 512)
    H, tn = o["time"][0], 0.37 * np.arange(40)
    s = o["arrivals"][0]
    np.testing.assert_allclose(H.sum(), s, rtol=1e-12)
    for k in (1, 2, 3):
        np.testing.assert_allclose((H * tn**k).sum() / s, T**k, rtol=2e-7)  # T from float32 tables
    assert np.all(o["late"] == 0) and np.all(o["early"] == 0)


def test_time_histogram_overflow_keeps_totals():
    """A grid ending before the arrivals puts them all in the late cell as (q, q T); one ending within two nodes of
    them splits them with signed weights, nodes plus overflow still summing to the arrivals."""
    a, n_idx = 1.5, 1.40
    T = a / (units.C_VACUUM / n_idx)
    src = L.sources.Point((0, 0, 0), lam=400.0, n_photons=1.0)
    det = _timed_sphere(0.0, 0.5, 8, a, n_idx)  # nodes 0 .. 3.5 ns, T = 7.0 ns
    o = O.light_of(*O.OracleRun(det).trace(src, 256, seed=3), 1, 1, 256)
    s = o["arrivals"][0]
    assert np.all(o["time"] == 0)
    np.testing.assert_allclose(o["late"][0], [s, s * T], rtol=2e-7)
    det = _timed_sphere(T - 1.75, 0.5, 4, a, n_idx)  # T half a node past the last node
    o = O.light_of(*O.OracleRun(det).trace(src, 256, seed=3), 1, 1, 256)
    tot = o["time"][0].sum() + o["late"][0, 0] + o["early"][0, 0]
    np.testing.assert_allclose(tot, s, rtol=1e-12)
    assert o["late"][0, 0] != 0


CE_KNOTS = np.array([0.0, 0.5, 1.2, 2.0, 2.6, np.pi])  # rad: uneven, both poles included
CE_VALUES = np.array([1.0, 0.9, 0.7, 0.5, 0.3, 0.2])


def _hat_means(knots):
    """E[hat_j(theta)] and E[hat_j(theta)^2] for theta of an isotropic direction (density sin(theta) / 2)."""
    th = np.linspace(0.0, np.pi, 200001)
    w = 0.5 * np.sin(th)
    m1, m2 = [], []
    for j in range(len(knots)):
        h = np.interp(th, knots, np.eye(len(knots))[j])
        m1.append(np.trapezoid(h * w, th))
        m2.append(np.trapezoid(h * h * w, th))
    return np.array(m1), np.array(m2)


def test_charge_cells_on_theta_knots_closed_form():
    """The charge histogram's theta axis (the cathode position angle): a centred isotropic source in an absorbing,
    detecting sensor sphere puts each photon's unit deposit at the polar angle of its direction, so cell j holds
    N E[hat_j(theta)] with theta of density sin(theta) / 2 (quadrature), and sum_j cell_j CE_j = N E[CE(theta)];
    the time histogram is weighted by CE(theta) (its total is N E[CE])."""
    from lumen_amd.params import Table
    from toys import sphere_toy

    det = sphere_toy(a=1.0, R=0.0, L_abs=None, ce=Table(CE_KNOTS, CE_VALUES)).timed(0.0, 0.5, 40)
    n = 1 << 17
    tal, h = O.OracleRun(det).trace(L.sources.Point((0, 0, 0), lam=400.0, n_photons=n), n, seed=21)
    light = O.light_of(tal, h, 1, len(CE_KNOTS), n)
    cells = light["charge"][0]
    m1, m2 = _hat_means(CE_KNOTS)
    z = (cells - n * m1) / np.sqrt(n * (m2 - m1**2))
    assert np.all(np.abs(z) < 5.0), z
    assert abs(cells.sum() - n) < 1e-6 * n  # the hats sum to one
    pe_ce = float(cells @ CE_VALUES)
    want = n * float(m1 @ CE_VALUES)
    assert abs(pe_ce - want) < 5.0 * np.sqrt(n * 0.1), (pe_ce, want)
    time_total = light["time"][0].sum() + light["late"][0, 0] + light["early"][0, 0]
    ce32 = CE_VALUES.astype(np.float32).astype(np.float64)  # the packed (float32) values the time weighting reads
    assert time_total == pytest.approx(float(cells @ ce32), rel=1e-9)
    # a wrong axis (the azimuth instead of the polar angle) fails the same check
    mz, _ = _hat_means(CE_KNOTS * 2.0)
    assert np.any(np.abs((cells - n * mz) / np.sqrt(n * (m2 + 1e-3))) > 10.0)


def test_diffusion_jump_is_the_plain_walk():
    """The diffusion jump (transport/fastpath.py) in a Rayleigh-dominated sphere (radius 2 m, Rayleigh 0.1 m,
    absorption 40 m): arrivals and mean arrival time with the jump agree with the plain walk's within their errors
    (the jump is exact), with far fewer geometry steps; the clearance grid never exceeds the distance to the
    spheres' surfaces."""
    from lumen_amd.detector.clearance import jump_grids
    from toys import sphere_toy

    det = sphere_toy(a=2.0, R=0.0, L_abs=40.0, L_ray=0.1)
    (g,) = jump_grids(det)
    ax = [g.lo[a] + g.h * np.arange(g.n[a]) for a in range(3)]
    X = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)
    r = np.linalg.norm(X, axis=1)
    exact = np.minimum(np.abs(r - 2.0), np.abs(r - 2.05))
    assert np.all(g.clear.reshape(-1) <= exact) and g.clear.max() > 1.5
    n, seeds = 1 << 12, 8
    src = L.sources.Point((0.1, -0.2, 0.05), lam=400.0, n_photons=1.0)
    o = O.OracleRun(det)
    res = {}
    for mode in (False, True):
        arr, tm, hs = [], [], []
        for k in range(seeds):  # independent replicas: their spread gives each mean's error
            tal, h = o.trace(src, n, seed=31 + k, controls=L.api.Controls(jump=mode))
            li = O.light_of(tal, h, 1, 1, n)
            arr.append(li["arrivals"][0])
            tm.append(li["mean_arrival_time"][0])
            hs.append(h)
        res[mode] = (np.array(arr), np.array(tm), np.sum(hs, 0))
    (ap, tp, hp), (aj, tj, hj) = res[False], res[True]
    for a, b in ((aj, ap), (tj, tp)):
        z = (a.mean() - b.mean()) / np.hypot(a.std(ddof=1), b.std(ddof=1)) * np.sqrt(seeds)
        assert abs(z) < 4.5, (a.mean(), b.mean(), z)
    assert 0.3 < ap.mean() < 0.9  # a diffusing, absorbing sphere: neither all nor none
    assert hj[12] < 0.2 * hp[12]  # vertices: the walk inside the balls makes no geometry step
    for h in (hp, hj):  # every photon's start weight ends in a fate
        assert abs(h[14] - sum(h[k] for k in (0, 1, 2, 3, 4, 5, 6, 7, 8))) < 1e-6 * h[14]


def _tori(tmp_path):
    from lumen_amd.detector.gdml import Recipe, load
    from toys import GDML_TORI

    f = tmp_path / "tori.gdml"
    f.write_text(GDML_TORI)
    return load(str(f), Recipe("tori", "World"))


def test_torus_closed_forms(tmp_path):
    """The oracle's torus roots where they are known: a ray across the ring's plane through its centre meets the
    tube at the centre +- (R +- r); a ray along the axis direction at radius R meets it at +- r."""
    det = _tori(tmp_path)
    names = [n.rsplit("/", 1)[-1] for n in det.nav.names]
    v1 = names.index("p1")
    o = O.OracleRun(det)
    cx, cy = -0.9, -0.9  # T1: R = 0.5, r = 0.2 about z
    starts = np.array([[cx - 1.0, cy, 0.0], [cx - 0.6, cy, 0.0], [cx - 0.1, cy, 0.0], [cx + 0.5, cy, -1.5]])
    dirs = np.array([[1.0, 0, 0], [1.0, 0, 0], [1.0, 0, 0], [0, 0, 1.0]])
    vol = o.locate(starts)
    assert vol.tolist() == [0, v1, 0, 0]
    t, va, nrm, _ = o.query(starts, dirs, vol)
    np.testing.assert_allclose(t, [0.3, 0.3, 0.4, 1.3], rtol=0, atol=1e-7)  # the tables store 0.2 and 0.9 in float32
    assert va.tolist() == [v1, 0, v1, v1]
    np.testing.assert_allclose(nrm, [[-1, 0, 0], [-1, 0, 0], [-1, 0, 0], [0, 0, -1]], atol=1e-6)


def test_torus_navigator_against_membership_march(tmp_path):
    """Random rays among the tori (one rotated, cut in phi, hollow): the oracle's next crossing is where the host
    solids' membership first changes along the ray (a march of 2 mm steps refined by bisection)."""
    from lumen_amd.detector.gdml import Recipe  # noqa: F401

    det = _tori(tmp_path)
    nav = det.nav
    o = O.OracleRun(det)
    rng = np.random.default_rng(8)
    centres = np.array([(-0.9, -0.9, 0.0), (0.9, 0.9, 0.0), (0.9, -0.9, 0.8)])
    x = centres[rng.integers(0, 3, 400)] + rng.uniform(-0.75, 0.75, (400, 3))
    d = rng.normal(size=(400, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    vol = o.locate(x)
    t, va, _, _ = o.query(x, d, vol)
    h = nav.host
    R, tt, kind = np.asarray(h["R"], np.float64), np.asarray(h["t"], np.float64), np.asarray(h["kind"])
    from lumen_amd.detector import solids  # noqa: F401

    def member(p):  # the navigator id holding each point (the deepest; the world's daughters do not overlap)
        out = np.full(len(p), nav.n, np.int64)
        for v in range(nav.n):  # the world (0) first, then its daughters
            xl = (p - tt[v]) @ R[v]
            out[_inside_kind(nav, int(kind[v]), xl)] = v
        return out

    steps = np.arange(1, 1501) * 0.002
    agree, earlier = 0, 0
    for i in range(len(x)):
        m = member(x[i] + steps[:, None] * d[i])
        ch = np.flatnonzero(m != vol[i])
        if not len(ch):
            assert t[i] > steps[-1] - 1e-9 or va[i] < 0
            agree += 1
            continue
        lo, hi = (steps[ch[0] - 1] if ch[0] else 0.0), steps[ch[0]]
        for _ in range(40):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if member((x[i] + mid * d[i])[None])[0] == vol[i] else (lo, mid)
        if abs(t[i] - hi) < 1e-7:  # a grazing root is found to sqrt(eps) of the tube's size
            agree += 1
        else:  # the march stepped over a sliver thinner than its step: the oracle's change lies before it
            earlier += int(t[i] < hi)
    assert agree >= 0.99 * len(x) and agree + earlier == len(x), (agree, earlier)


def _inside_kind(nav, k, xl):
    """Membership of points (kind frame) by the CSG tables' leaves and program, on the host (float64)."""
    from test_host import _inside_kind as ik

    return ik(nav, k, xl)


def _ray_hits_tube(o, d, R, t, radius, half):
    """Whether rays from o along d (n, 3) meet a closed tube (radius, half-length about its local z axis, placed at
    x_w = R x + t): the slab of its caps intersected with the infinite cylinder's interval (closed form)."""
    ol, dl = (o - t) @ R, d @ R
    a = dl[:, 0] ** 2 + dl[:, 1] ** 2
    b = 2 * (ol[0] * dl[:, 0] + ol[1] * dl[:, 1])
    c = ol[0] ** 2 + ol[1] ** 2 - radius**2
    disc = b * b - 4 * a * c
    ok = disc >= 0
    sq = np.sqrt(np.maximum(disc, 0.0))
    with np.errstate(divide="ignore", invalid="ignore"):
        c0, c1 = (-b - sq) / (2 * a), (-b + sq) / (2 * a)
        z0, z1 = (-half - ol[2]) / dl[:, 2], (half - ol[2]) / dl[:, 2]
    s0, s1 = np.minimum(z0, z1), np.maximum(z0, z1)
    lo, hi = np.maximum(np.maximum(c0, s0), 0.0), np.minimum(c1, s1)
    return ok & (a > 0) & (lo <= hi)


def test_first_step_records_are_the_pmts_solid_angle():
    """The tutorial run cut after one vertex: a record at every first step whose ray enters a PMT, so the count is
    binomial in 2^18 with the PMTs' share of the directions from the source, found here by closed-form ray-tube
    intersections over 4e6 random directions (the 8 discs: r 50 mm, 20 mm thick)."""
    det = example_tank()
    o = O.OracleRun(det)
    n = 1 << 18
    src = L.sources.Point((0.0, 0.0, 0.1), lam=400.0, n_photons=2e5)
    _, h = o.trace(src, n, seed=0, controls=L.api.Controls(max_vertices=1))
    rng = np.random.default_rng(1)
    d = rng.normal(size=(4_000_000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x0 = np.array([0.0, 0.0, 0.1])
    nav = det.nav
    hit = np.zeros(len(d), bool)
    for v in range(nav.n):
        if det.sensor_of_vol[v] >= 0:
            hit |= _ray_hits_tube(x0, d, np.asarray(nav.host["R"][v], np.float64),
                                  np.asarray(nav.host["t"][v], np.float64), 0.05, 0.01)
    p = hit.mean()
    rec = h[L._lib.HEALTH_NAMES.index("records")]
    assert abs(rec - n * p) < 4.0 * math.sqrt(n * p * (1 - p)) + 4.0 * n * math.sqrt(p / len(d)), (rec, n * p)
