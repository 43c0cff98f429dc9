"""The oracle's Mie scattering (G4OpMieHG, the reference's media/phase.py:98-190) pinned without the engine.

- The drawn scattering angle against the mixture's closed-form CDF in 40 equal-mass bins (chi2), every sign and
  size of g and the ratio's edges (tests/unit/test_mie.py:158-175).
- The Henyey-Greenstein moments <P_l(mu)> = r g_f^l + (1 - r)(-g_b)^l, l = 1..4 (tests/unit/test_mie.py:211-221).
- Geant4 10.7.4 / 11.2.2 single-scatter cos(theta) histograms of the g4ref Mie toy (tests/golden/g4_mie*_hist.npz,
  the reference's tests/data/g4ref_mie fixtures; the two versions' files are identical), with the two wrong
  readings the fixtures carry (no mirror of the backward lobe; the ratio swapped) failing.
- The density: normalised, the draw's own density in bins, and the live/drawn weight unbiased (mean 1, and the
  live first moment), also at the ratio's edges where the draw takes the defensive cover.
- The whole loop on the Mie sphere: every photon's expected deposits sum to 1 (no absorption), and the direct light
  (the first vertex's deposit, arriving at R n / c) is exp(-R / L) per photon.
"""

import json
import math
from pathlib import Path

import numpy as np
import pytest
from scipy import stats

import lumen_amd as L
from oracle import oracle as O
from toys import MIE_CASES, MIE_L, MIE_N, MIE_R, mie_cdf_angle, mie_toy

GOLDEN = Path(__file__).parent / "golden"
P_MIN = 1e-4
G = (0.0, 0.3, 0.9, 0.999)
CASES = [(gf, gb, r) for gf in G for gb in G for r in (0.0, 0.3, 1.0)]
CASES += [(-0.5, 0.4, 0.6), (0.7, -0.6, 0.2), (-0.999, -0.999, 0.5), (0.6, 0.3, 0.7), (0.0, 0.8, 0.5)]
WI = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])


def angle(wo, wi):
    """The angle between unit vectors, accurate at both poles: 2 atan2(|wo - wi|, |wo + wi|)."""
    return 2.0 * np.arctan2(np.linalg.norm(wo - wi, axis=-1), np.linalg.norm(wo + wi, axis=-1))


def cover(r):
    return 0.01 if r <= 0.0 else 0.99 if r >= 1.0 else r


@pytest.mark.parametrize("case", CASES)
def test_angle_follows_the_mixture_in_bins_of_exact_mass(case):
    n = 200_000
    gf, gb, r = case
    rc = cover(r)  # a ratio of 0 or 1 is drawn at its defensive cover
    wo, _ = O.mie(WI, gf, gb, r, n, seed=11)
    th = angle(wo, WI)
    grid = np.unique(np.concatenate([[0.0], np.geomspace(1e-8, 0.5 * np.pi, 20000),
                                     np.pi - np.geomspace(1e-8, 0.5 * np.pi, 20000), [np.pi]]))
    edges = np.interp(np.linspace(0, 1, 41), mie_cdf_angle(grid, gf, gb, rc), grid)
    edges[0], edges[-1] = 0.0, np.pi
    obs, _ = np.histogram(th, edges)
    exp = np.diff(mie_cdf_angle(edges, gf, gb, rc)) * n
    ok = exp > 20
    c2 = float(np.sum((obs[ok] - exp[ok]) ** 2 / exp[ok]))
    assert stats.chi2.sf(c2, int(ok.sum()) - 1) > P_MIN, (case, c2, int(ok.sum()))
    assert np.allclose(np.linalg.norm(wo, axis=1), 1.0, atol=1e-12)


@pytest.mark.parametrize("case", [(0.0, 0.0, 0.5), (0.6, 0.3, 0.7), (0.0, 0.8, 0.5), (0.95, 0.0, 0.99),
                                  (0.999, 0.999, 0.3), (-0.5, 0.4, 0.6)])
def test_henyey_greenstein_moments(case):
    gf, gb, r = case
    wo, _ = O.mie(WI, gf, gb, r, 1_000_000, seed=17)
    mu = wo @ WI
    for order in range(1, 5):
        pl = np.polynomial.legendre.legval(mu, [0] * order + [1])
        truth = r * gf**order + (1 - r) * (-gb) ** order
        z = (pl.mean() - truth) / (pl.std(ddof=1) / math.sqrt(len(pl)))
        assert abs(z) < 4.5, (case, order, pl.mean(), truth, z)


@pytest.mark.parametrize("name", ["mieA", "mieB"])
def test_single_scatter_angles_match_geant4(name):
    man = json.loads((GOLDEN / "g4_mie_hist_manifest.json").read_text())
    case = tuple(man["cases"][name][k] for k in ("g_forward", "g_backward", "forward_ratio"))
    with np.load(GOLDEN / f"g4_{name}_hist.npz") as z:
        edges, obs, err = z["edges"], z["obs"], z["err"]
        exact, no_mirror, swapped = z["A_mirror"], z["B_nomirror"], z["C_swapratio"]
    n = 4_000_000
    wo, _ = O.mie(WI, *case, n, seed=23)
    h, _ = np.histogram(wo @ WI, edges)
    ours, ours_err = h / n, np.sqrt(h) / n
    c2 = float(np.sum((obs - ours) ** 2 / (err**2 + ours_err**2)))
    assert stats.chi2.sf(c2, len(obs) - 1) > P_MIN, (name, c2)
    c2_exact = float(np.sum((ours - exact) ** 2 / np.maximum(exact / n, 1e-30)))
    assert stats.chi2.sf(c2_exact, len(obs) - 1) > P_MIN, (name, c2_exact)
    for wrong in (no_mirror, swapped):  # the readings Geant4's histogram rules out are ruled out here too
        if np.allclose(wrong, exact, rtol=0, atol=1e-9):  # a ratio of 1/2 swaps into itself
            continue
        c2w = float(np.sum((ours - wrong) ** 2 / np.maximum(wrong / n, 1e-30)))
        assert stats.chi2.sf(c2w, len(obs) - 1) < 1e-12, (name, c2w)


@pytest.mark.parametrize("case", [(0.6, 0.3, 0.7), (0.999, 0.0, 0.5), (-0.5, 0.4, 0.6), (0.9, 0.3, 1.0),
                                  (0.3, 0.9, 0.0)])
def test_the_density_is_normalised_and_the_draws_own(case):
    gf, gb, r = case
    th = np.unique(np.concatenate([np.geomspace(1e-9, 0.5 * np.pi, 40000), np.pi - np.geomspace(1e-9, 0.5 * np.pi,
                                                                                                  40000)]))
    f = O.mie_pdf(gf, gb, r, np.cos(th)) * np.sin(th)
    assert abs(np.sum(0.5 * (f[1:] + f[:-1]) * np.diff(th)) * 2 * np.pi - 1.0) < 1e-4
    # the density times the solid angle of each bin is the CDF difference
    edges = np.linspace(0.0, np.pi, 9)
    for a, b in zip(edges[:-1], edges[1:]):
        t = np.linspace(a, b, 20001)
        g = O.mie_pdf(gf, gb, r, np.cos(t)) * np.sin(t) * 2 * np.pi
        mass = np.sum(0.5 * (g[1:] + g[:-1]) * np.diff(t))
        want = mie_cdf_angle(b, gf, gb, r) - mie_cdf_angle(a, gf, gb, r)
        assert mass == pytest.approx(want, rel=1e-3, abs=1e-9), (case, a, b)


@pytest.mark.parametrize("drawn,live", [((0.6, 0.3, 0.7), (0.5, 0.35, 0.75)), ((0.95, 0.3, 1.0), (0.95, 0.3, 0.9)),
                                        ((0.0, 0.8, 0.0), (0.2, 0.8, 0.3)), ((0.6, 0.3, 0.7), (0.6, 0.3, 0.7))])
def test_live_over_drawn_weights_are_unbiased(drawn, live):
    n = 1_000_000
    wo, w = O.mie(WI, *drawn, n, seed=29, live=live)
    mu = wo @ WI
    z = (w.mean() - 1.0) / (w.std(ddof=1) / math.sqrt(n) + 1e-300)
    assert abs(z) < 4.5 or (w.std() == 0 and w.mean() == pytest.approx(1.0, abs=1e-12)), (drawn, live, w.mean())
    gf, gb, r = live
    m1 = r * gf - (1 - r) * gb  # the live mixture's mean cosine
    zz = ((w * mu).mean() - m1) / ((w * mu).std(ddof=1) / math.sqrt(n))
    assert abs(zz) < 4.5, (drawn, live, (w * mu).mean(), m1)
    if drawn == live and 0.0 < drawn[2] < 1.0:  # drawn at the live values, away from the edges: exactly 1
        assert np.all(np.abs(w - 1.0) < 1e-12)


@pytest.mark.parametrize("case", ["A", "water"])
def test_mie_sphere_charge_and_direct_light(case):
    """Every photon's expected deposits sum to 1 in expectation; the first vertex deposits exp(-R / L) at R n / c
    (the run cut after one vertex: every photon deposits exactly that)."""  # noqa: D205, D209
    det = mie_toy(case, 1)
    n = 1 << 15
    src = L.sources.Point((0.0, 0.0, 0.0), lam=L.units.HC / 3.0, n_photons=n)
    o = O.OracleRun(det)
    tal, h = o.trace(src, n, seed=3)
    light = O.light_of(tal, h, 1, 1, n)
    q, err = float(light["arrivals"][0]), float(light["arrivals_error"][0])
    assert abs(q - n) < 5 * err and err > 0, (case, q, err)
    tal1, h1 = o.trace(src, n, seed=3, controls=L.api.Controls(max_vertices=1))
    one = O.light_of(tal1, h1, 1, 1, n)
    assert float(one["arrivals"][0]) == pytest.approx(n * math.exp(-MIE_R / MIE_L), rel=1e-6), case
    assert float(one["mean_arrival_time"][0]) == pytest.approx(MIE_R * MIE_N / L.units.C_VACUUM, rel=1e-6)
    assert MIE_CASES[case]
