"""GLISUR's facet paths on the device, in the ring counter (a beam onto a flat GLISUR interface inside 36 polar rings
of 5 degrees): the rings against Geant4 11.3.0's (the reference's tests/data/g4_boundary fixtures, the two-sample
chi2 rule of tests/unit/test_g4_boundary.py:35-77), the same photons as the oracle's, and the adjoint in polish
against common-random-number central differences."""

import json
import os

import numpy as np
import pytest
import torch
from scipy import stats

import lumen_amd as L
from lumen_amd.autograd import simulate_diff, table_tensor
from oracle import oracle as O
from toys import glisur_beam, glisur_toy

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
MANIFEST = json.load(open(os.path.join(GOLDEN, "g4_glisur_manifest.json")))["cases"]
RANDOM_CASES = sorted(k for k, v in MANIFEST.items() if v["pol"] == "random")
KEY_P = "surfaces/Glisur/polish"


def _chi2_p(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    n1, n2 = a.sum(), b.sum()
    use = (a + b) > 0
    c2 = float(np.sum((a[use] / n1 - b[use] / n2) ** 2 / ((a[use] + b[use]) / (n1 * n2))))
    return float(stats.chi2.sf(c2, int(use.sum()) - 1))


@pytest.mark.parametrize("name", RANDOM_CASES)
def test_device_rings_are_geant4s(gpu, name):
    print(f"[glisur] {name}", flush=True)
    case = MANIFEST[name]
    det = glisur_toy("dielectric", case["n1"], case["n2"], case["value"])
    n = 1 << 20
    g = L.simulate(det, glisur_beam(case["theta"]), samples=n, seed=case["seed"])
    got = g.arrivals.cpu().numpy() * n
    assert abs(got.sum() - n) < 1e-4 * n
    with np.load(os.path.join(GOLDEN, case["file"])) as z:
        assert _chi2_p(z["rings"], got) > 1e-3, name
    # the same photons as the oracle's: a photon changes ring only within float32 rounding of a ring's edge
    m = 1 << 16
    gd = L.simulate(det, glisur_beam(case["theta"]), samples=m, seed=3).arrivals.cpu().numpy()
    tal, h = O.OracleRun(det).trace(glisur_beam(case["theta"]), m, seed=3)
    o = O.light_of(tal, h, 36, 1, m)["arrivals"]
    assert np.abs(gd - o).sum() < 2e-3, np.abs(gd - o).sum()


@pytest.mark.parametrize("kind,keyed", [("metal", False), ("metal", True), ("dielectric", True)])
def test_device_same_photons_as_oracle(gpu, kind, keyed):
    """Metal and dielectric paths, plain and under the cover mixture (polish a key), with absorption and detection
    on the GLISUR row: the device's rings and health against the oracle's photons."""
    print(f"[glisur] {kind} keyed={keyed}", flush=True)
    det = glisur_toy(kind, 1.0, 1.5, 0.3, R=0.8, T=0.1 if kind == "dielectric" else 0.0, eff=0.5, keyed=keyed)
    m = 1 << 16
    src = glisur_beam(57.5)  # the straight share at polar 122.5 degrees: mid-ring, not on a ring's edge
    g = L.simulate(det, src, samples=m, seed=8)
    tal, h = O.OracleRun(det).trace(src, m, seed=8)
    o = O.light_of(tal, h, 36, 1, m)["arrivals"]
    a = g.arrivals.cpu().numpy()
    assert np.abs(a - o).sum() < 2e-3 * max(o.sum(), 1e-3), (np.abs(a - o).sum(), o.sum())
    assert abs(g.health.balance()) < 1e-5 * g.health.start_weight


@pytest.mark.parametrize("kind", ["metal", "dielectric"])
def test_polish_adjoint_matches_crn_finite_differences(gpu, kind):
    """d/d polish of a ring-weighted count: the adjoint against one-sided differences at a fixed seed with every
    random choice drawn at the unperturbed polish (the facets from the cover mixture).  One-sided toward a higher
    polish: the live ball (radius 1 - polish) then lies inside the drawn one, so the reweighted estimate is unbiased
    (a smaller polish widens the live support past what was drawn).  p_m vanishes like q ~ sqrt(dr) at the cone's
    edge, so the light is C^1 in r with a second derivative that diverges there: a difference quotient over a step
    h carries a sqrt(h) term, cancelled by Richardson's extrapolation over h and h / 4, 2 D(h / 4) - D(h)."""
    print(f"[glisur] adjoint {kind}", flush=True)
    det = glisur_toy(kind, 1.0, 1.5, 0.3, keyed=True)
    src = glisur_beam(57.5)
    p0 = det.params()
    c = torch.cos(torch.deg2rad(torch.arange(36, dtype=torch.float64) * 5.0 + 2.5)).cuda() ** 2
    n, seed = 1 << 20, 11
    t = table_tensor(p0, KEY_P)
    light = simulate_diff(det, src, {KEY_P: t}, samples=n, seed=seed)
    (light.charge[..., 0].to(torch.float64) * c).sum().backward()
    g = float(t.grad.sum())
    h, v = 4e-5, float(p0[KEY_P])

    def f(s):
        lt = L.simulate(det, src, p0.set(KEY_P, v * s), samples=n, seed=seed, reference=p0)
        return float((lt.charge[..., 0].to(torch.float64) * c).sum())

    f0 = f(1.0)
    d1, d4 = (f(1 + h) - f0) / (h * v), (f(1 + h / 4) - f0) / (h / 4 * v)
    fd = 2.0 * d4 - d1
    assert abs(g - fd) <= 1e-2 * abs(fd) + 1e-5, (g, fd, d1, d4)
    assert abs(fd) > 1e-3  # polish moves the rings
