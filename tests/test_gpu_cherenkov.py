"""The device's Cherenkov emission on the cone toy (a track along +z in plain water inside a ring counter): the
rings' shares of the Frank-Tamm yield and their mean arrival times against the geometry's closed form, and the same
photons as the oracle's (the reference's ``sources/cherenkov.py:33-66``)."""

import numpy as np
import pytest

import lumen_amd as L
from oracle import oracle as O
from toys import cone_rings_closed_form, cone_toy

pytestmark = pytest.mark.gpu

RINGS, BETA, N_IDX = 90, 0.9, 1.33


@pytest.fixture(scope="module")
def det():
    return cone_toy(RINGS, N_IDX)


def _src(beta=BETA):
    return L.sources.CherenkovSegment("Water", start=(0.0, 0.0, -0.15), direction=(0.0, 0.0, 1.0), length=0.3,
                                      beta=beta)


@pytest.mark.parametrize("beta", [BETA, 1.0])
def test_cone_toy_closed_form_on_device(gpu, det, beta):
    n = 1 << 20
    g = L.simulate(det, _src(beta), samples=n, seed=4)
    a, e = g.arrivals.cpu().numpy(), g.arrivals_error.cpu().numpy()
    total = g.health.start_weight
    share, tmean = cone_rings_closed_form(RINGS, -0.15, 0.15, 1.0 / (np.float32(beta) * N_IDX), np.float32(beta),
                                          N_IDX)
    hit = share > 0
    assert abs(a.sum() / total - 1.0) < 1e-5  # every photon reaches a ring
    np.testing.assert_allclose(a[~hit], 0.0, atol=0.0)
    pulls = (a[hit] / total - share[hit]) / np.maximum(e[hit] / total, 1e-12)
    assert np.abs(pulls).max() < 5.0, pulls
    big = share > 0.05
    np.testing.assert_allclose(g.mean_arrival_time.cpu().numpy()[big], tmean[big], rtol=1e-3)


def test_cone_toy_same_photons_as_oracle(gpu, det):
    n = 1 << 18
    g = L.simulate(det, _src(), samples=n, seed=6)
    tal, h = O.OracleRun(det).trace(_src(), n, seed=6)
    o = O.light_of(tal, h, RINGS, 1, n)
    a = g.arrivals.cpu().numpy()
    # the same photons: a photon lands in another ring only when it crosses within float32 rounding of a ring's edge
    np.testing.assert_allclose(a, o["arrivals"], rtol=2e-3, atol=2e-3 * float(a.max()))
    assert abs(g.health.start_weight - float(h[14])) < 1e-6 * float(h[14])
    np.testing.assert_allclose(g.mean_arrival_time.cpu().numpy()[a > 0.01 * a.max()],
                               o["mean_arrival_time"][a > 0.01 * a.max()], rtol=1e-5)
