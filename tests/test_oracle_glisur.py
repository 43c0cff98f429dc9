"""GLISUR's facet paths in the oracle (the reference's surfaces/facets.py GLISUR parts, Geant4's GetFacetNormal with
the glisur model and the DielectricMetal / DielectricDielectric facet loops):

- the facet sampler against a brute-force draw of points uniform in the ball B(n, r) (two-sample chi2 on the facet's
  cosine), and its log density against the ray-through-a-ball integral p_m = 2 q (3 c^2 + q^2) / (3 Vol) (normalised);
- the metal path at normal incidence: one facet with tilt below 45 degrees always leaves, so the last direction's
  cosine is 2 c^2 - 1 of the facet's (chi2 against p_m mapped);
- the dielectric path against Geant4 11.3.0's histograms (tests/golden/g4_glisur_*.npz, the reference's fixtures: six
  cases with random linear states, two with s or p states compared under Geant4's conventions for the state between
  facets), the two-sample chi2 on the (dz, |phi|) histogram of the last direction (the reference's rule,
  tests/unit/test_g4_boundary.py:35-77)."""

import json
import math
import os

import numpy as np
import pytest
from scipy import integrate, stats

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
MANIFEST = json.load(open(os.path.join(GOLDEN, "g4_glisur_manifest.json")))["cases"]
P_MIN = 1e-3


def _chi2_p(a, b):
    a, b = np.ravel(a).astype(np.float64), np.ravel(b).astype(np.float64)
    n1, n2 = a.sum(), b.sum()
    use = (a + b) > 0
    c2 = float(np.sum((a[use] / n1 - b[use] / n2) ** 2 / ((a[use] + b[use]) / (n1 * n2))))
    return float(stats.chi2.sf(c2, int(use.sum()) - 1))


def _pm(c, r):
    """p_m per solid angle at facet cosine c for V along the normal (no cut: the whole ball)."""
    q2 = r * r - (1.0 - c * c)
    q = np.sqrt(np.maximum(q2, 0.0))
    return np.where(q2 > 0, 2.0 * q * (3.0 * c * c + q2) / (3.0 * (4.0 / 3.0 * math.pi * r**3)), 0.0)


@pytest.mark.parametrize("r", [0.05, 0.3, 0.7, 0.99])
def test_glisur_facets_are_directions_to_points_in_the_ball(r):
    n = 200_000
    m, lp = O.glisur_facets(r, [0.0, 0.0, 1.0], n, seed=3)
    rng = np.random.default_rng(4)
    s = rng.normal(size=(n, 3))
    s *= (rng.uniform(size=n) ** (1.0 / 3.0) / np.linalg.norm(s, axis=1))[:, None]
    M = np.array([0.0, 0.0, 1.0]) + r * s
    c_ref = M[:, 2] / np.linalg.norm(M, axis=1)
    edges = np.linspace(math.sqrt(max(1.0 - r * r, 0.0)), 1.0, 31)
    assert _chi2_p(np.histogram(m[:, 2], edges)[0], np.histogram(c_ref, edges)[0]) > P_MIN
    np.testing.assert_allclose(np.linalg.norm(m, axis=1), 1.0, atol=1e-12)
    np.testing.assert_allclose(lp, np.log(_pm(m[:, 2], r)), rtol=1e-10, atol=1e-12)  # the ray-through-a-ball integral
    lo = math.sqrt(max(1.0 - r * r, 0.0))
    assert abs(integrate.quad(lambda c: 2 * math.pi * _pm(np.array(c), r), lo, 1.0, limit=200)[0] - 1.0) < 1e-7


def test_glisur_cover_mixture_density():
    """Drawn from the cover (1 - e) p_m + e p_shell with e = 0.3, the weights p_m / q average 1 (the mixture is a
    density whose support holds p_m's)."""
    m, lq = O.glisur_facets(0.4, [0.3, 0.1, 0.95], 200_000, seed=5, e=0.3)
    _, lp = O.glisur_facets(0.4, [0.3, 0.1, 0.95], 1, seed=5)
    V = np.array([0.3, 0.1, 0.95]) / np.linalg.norm([0.3, 0.1, 0.95])
    keep = (m @ V) > 0
    assert keep.all()
    c, q2 = m[:, 2], 0.16 - (m[:, 0] ** 2 + m[:, 1] ** 2)
    d = max(0.4 - V[2], 0.0)
    vol = 4 / 3 * math.pi * 0.4**3 - math.pi * d * d * (0.8 + V[2]) / 3
    pm = np.where(q2 > 0, 2 * np.sqrt(np.maximum(q2, 0)) * (3 * c * c + q2) / (3 * vol), 0.0)
    w = pm / np.exp(lq)
    assert abs(w.mean() - 1.0) < 4 * w.std() / math.sqrt(len(w)) + 1e-3
    assert w.max() <= 1.0 / 0.7 + 1e-9  # p_m / q <= 1 / (1 - e)


def test_glisur_metal_path_at_normal_incidence():
    r = 0.5
    wo, far = O.glisur_paths(r, [0.0, 0.0, -1.0], 200_000, seed=6)
    assert not far.any() and (wo[:, 2] > 0).all()
    cmin = math.sqrt(1 - r * r)
    edges_c = np.linspace(cmin, 1.0, 25)
    mass = np.array([integrate.quad(lambda c: 2 * math.pi * _pm(np.array(c), r), a, b)[0]
                     for a, b in zip(edges_c[:-1], edges_c[1:])])
    got = np.histogram(np.sqrt((wo[:, 2] + 1.0) / 2.0), edges_c)[0]  # wo_z = 2 c^2 - 1
    chi2 = float((((got - mass * len(wo)) ** 2) / (mass * len(wo))).sum())
    assert stats.chi2.sf(chi2, len(mass) - 1) > P_MIN


@pytest.mark.parametrize("name", sorted(MANIFEST))
def test_glisur_dielectric_path_is_geant4s(name):
    case = MANIFEST[name]
    theta = math.radians(case["theta"])
    wi = np.array([math.sin(theta), 0.0, -math.cos(theta)])
    n = 1 << 18
    s = np.array([0.0, 1.0, 0.0])
    p = np.cross(wi, s)
    if case["pol"] == "random":
        phi = np.random.default_rng(case["seed"]).uniform(0.0, 2.0 * math.pi, n)[:, None]
        er = np.cos(phi) * p + np.sin(phi) * s
    else:
        er = np.tile(s if case["pol"] == "s" else p, (n, 1))
    wo, far = O.glisur_paths(1.0 - case["value"], wi, n, seed=case["seed"], er=er, dielectric=True, n1=case["n1"],
                             n2=case["n2"], g4conv=case["pol"] != "random")
    assert np.array_equal(far, wo[:, 2] < 0.0)
    with np.load(os.path.join(GOLDEN, case["file"])) as z:
        mine = np.histogram2d(wo[:, 2], np.abs(np.arctan2(wo[:, 1], wo[:, 0])), bins=[z["edges_z"], z["edges_phi"]])[0]
        assert _chi2_p(z["hist"], mine) > P_MIN, name


RANDOM_CASES = sorted(k for k, v in MANIFEST.items() if v["pol"] == "random")


def _ring_counts(tal, h, n_photons):
    """The 36 rings' counts from a trace's arrivals (each photon of weight 1/n ends in one ring)."""
    return O.light_of(tal, h, 36, 1, n_photons)["arrivals"] * n_photons


@pytest.mark.parametrize("name", RANDOM_CASES)
def test_glisur_toy_in_the_oracle_counts_geant4s_rings(name):
    """The oracle's trace through the ring counter (the path within a detector: the side it ends on decides the
    medium it goes on in) against Geant4's ring histogram."""
    from toys import glisur_beam, glisur_toy

    case = MANIFEST[name]
    det = glisur_toy("dielectric", case["n1"], case["n2"], case["value"])
    n = 1 << 16
    tal, h = O.OracleRun(det).trace(glisur_beam(case["theta"]), n, seed=case["seed"])
    got = _ring_counts(tal, h, n)
    assert abs(got.sum() - n) < 1e-6 * n  # every photon counted once
    with np.load(os.path.join(GOLDEN, case["file"])) as z:
        assert _chi2_p(z["rings"], got) > P_MIN, name


@pytest.mark.parametrize("keyed", [False, True])
def test_glisur_metal_toy_in_the_oracle(keyed):
    """GLISUR_METAL in the ring counter: the rings of the oracle's trace against its facet path's own directions;
    keyed, drawn from the cover mixture, the weights average out to the same rings."""
    from toys import glisur_beam, glisur_toy

    det = glisur_toy("metal", polish=0.4, keyed=keyed)
    n = 1 << 16
    tal, h = O.OracleRun(det).trace(glisur_beam(40.0), n, seed=9)
    got = _ring_counts(tal, h, n)
    t = np.deg2rad(40.0)
    wo, _ = O.glisur_paths(0.6, [np.sin(t), 0.0, -np.cos(t)], 1 << 18, seed=10)
    ref = np.histogram(np.degrees(np.arccos(np.clip(wo[:, 2], -1, 1))), np.linspace(0, 180, 37))[0]
    assert abs(got.sum() - n) < 1e-3 * n
    assert got[18:].sum() == 0.0
    if not keyed:
        assert _chi2_p(ref, got) > P_MIN
    else:  # weighted counts: compare shares within their own errors
        share = ref / ref.sum()
        e = np.sqrt(share * (1 - share) / ref.sum() + share / n * 1.2)
        assert np.all(np.abs(got / n - share) < 5 * e + 1e-4)


def test_float32_model_keeps_the_crossed_face_behind_a_crossing_near_the_origin():
    """A crossing at the origin reached from 0.5 m away rounds off the face by up to its ray's band (~1e-7 m), far
    above the band of the crossing point's own coordinates (the 1e-9 m floor): the float32 model (the engine's
    arithmetic and band rule) takes the face behind by the crossing ray's scale, so it walks the same paths as FP64,
    with no zero-length steps and no pushes."""
    from toys import glisur_beam, glisur_toy

    det = glisur_toy("dielectric", 1.67, 1.8, 0.01)
    n = 1 << 14
    a, b = (O.OracleRun(det).trace(glisur_beam(45.0), n, seed=3, fp32=f) for f in (False, True))
    ra, rb = (O.light_of(t, h, 36, 1, n)["arrivals"] for t, h in (a, b))
    assert np.abs(ra - rb).sum() < 2e-3
    assert float(b[1][9]) == 0.0 and float(b[1][12]) == float(a[1][12])  # pushed; vertices
