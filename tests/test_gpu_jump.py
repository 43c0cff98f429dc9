"""The diffusion jump (the reference's transport/fastpath.py) on the device, through the C ABI: against the oracle's
jump on the same photons, against the engine's own plain walk (seed replicas), and its adjoint terms against
common-random-number finite differences."""

import numpy as np
import pytest

import lumen_amd as L
from lumen_amd.autograd import simulate_diff, table_tensor
from oracle import oracle as O
from toys import sphere_toy

pytestmark = pytest.mark.gpu

JUMP = L.api.Controls(jump=True)
SRC = L.sources.Point((0.1, -0.2, 0.05), lam=400.0, n_photons=1.0)


@pytest.fixture(scope="module")
def det():
    return sphere_toy(a=2.0, R=0.0, L_abs=40.0, L_ray=0.1)


def test_jump_same_photons_as_oracle(gpu, det):
    n = 1 << 14
    g = L.simulate(det, SRC, samples=n, seed=41, controls=JUMP)
    tal, h = O.OracleRun(det).trace(SRC, n, seed=41, controls=JUMP)
    o = O.light_of(tal, h, 1, 1, n)
    z = (float(g.arrivals[0]) - o["arrivals"][0]) / np.hypot(float(g.arrivals_error[0]), o["arrivals_error"][0])
    assert abs(z) < 4.5, (float(g.arrivals[0]), o["arrivals"][0])
    assert abs(float(g.mean_arrival_time[0]) / o["mean_arrival_time"][0] - 1.0) < 0.02
    assert g.health.vertices < 1.5 * h[12] and g.health.vertices > 0.5 * h[12]
    assert abs(g.health.balance()) < 1e-4 * g.health.start_weight


def test_jump_equals_the_plain_walk(gpu, det):
    seeds, n = 8, 1 << 16
    out = {}
    for mode in (False, True):
        lights = [L.simulate(det, SRC, samples=n, seed=50 + k, controls=L.api.Controls(jump=mode))
                  for k in range(seeds)]
        out[mode] = (np.array([float(x.arrivals[0]) for x in lights]),
                     np.array([float(x.mean_arrival_time[0]) for x in lights]),
                     sum(x.health.vertices for x in lights))
    (ap, tp, vp), (aj, tj, vj) = out[False], out[True]
    for a, b in ((aj, ap), (tj, tp)):
        z = (a.mean() - b.mean()) / np.hypot(a.std(ddof=1), b.std(ddof=1)) * np.sqrt(seeds)
        assert abs(z) < 4.5, (a.mean(), b.mean(), z)
    assert vj < 0.2 * vp


@pytest.mark.parametrize("key", ["materials/Water/ABSLENGTH", "materials/Water/RAYLEIGH"])
def test_jump_adjoint_against_crn_finite_differences(gpu, det, key):
    """d(arrivals)/d(scale of the table) with the jump: the walk's terms -L d mu_abs and n d mu / mu - L d mu."""
    p0 = det.params()
    n = 1 << 18
    t = table_tensor(p0, key)
    light = simulate_diff(det, SRC, {key: t}, samples=n, seed=61, controls=JUMP)
    light.arrivals.sum().backward()
    base = np.asarray(p0[key].value, np.float64)
    g = float((t.grad.cpu().numpy() * base).sum())

    def tot(scale):
        p = p0.set(key, L.params.Table(p0[key].energy, base * scale))
        return float(L.simulate(det, SRC, p, samples=n, seed=61, reference=p0, controls=JUMP).arrivals.sum())

    h = 1e-3
    fd = (tot(1 + h) - tot(1 - h)) / (2 * h)
    assert abs(g - fd) <= 0.02 * abs(fd) + 1e-4 and abs(fd) > 1e-2, (key, g, fd)
