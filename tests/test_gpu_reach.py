"""The reach bound (the reference's geometry/reach.py:1-40, 174-222: a step whose free path ends inside its volume's
box, with no sensor volume's box on the ray before that box's exit, queries only a few bands past the free path)
changes no photon: with it and without it the same vertices, fates, tallies and gradients, on the LAr box (where
most steps are bounded), the water tank, the navigator's edge-case scenes and a batched LAr launch."""

import numpy as np
import pytest

import lumen_amd as L
import nav_scenes as S
from lumen_amd.api import Controls, engine
from lumen_amd.autograd import simulate_diff, table_tensor
from lumen_amd.detector import lar_box, muon_track, voxel_points, water_tank
from lumen_amd.sources import Batch

pytestmark = pytest.mark.gpu
ON = Controls(reach_bound=True)


@pytest.fixture(scope="module")
def lar():
    return lar_box()


def _same(a, b, what):
    ha, hb = np.asarray(a.health), np.asarray(b.health)
    # the step counts are exact: every photon takes the same steps
    assert ha[12] == hb[12] and ha[13] == hb[13] and ha[10] == hb[10], (what, ha, hb)
    # the fates are float32 per-lane sums added in any order
    np.testing.assert_allclose(ha[:15], hb[:15], rtol=1e-6, atol=1e-9, err_msg=what)
    for k in ("arrivals", "arrival2", "charge"):
        np.testing.assert_allclose(getattr(a, k).cpu().numpy(), getattr(b, k).cpu().numpy(), rtol=1e-12, atol=0,
                                   err_msg=f"{what}: {k}")


def test_lar_box_the_same_with_and_without_the_bound(gpu, lar):
    pts = voxel_points(n=2)
    batch = Batch(tuple(pts), tuple([1 << 14] * len(pts)))
    a = L.simulate(lar, batch, seed=5, controls=ON)
    assert engine(lar).last_launch()["features"] & 32
    b = L.simulate(lar, batch, seed=5)
    assert not engine(lar).last_launch()["features"] & 32
    _same(a, b, "lar")
    print(f"[reach] lar: {a.health.bounded / a.health.vertices:.3f} of the steps bounded", flush=True)
    assert a.health.bounded > 0.0 and b.health.bounded == 0.0
    # most of a long run's steps are the shifted light's reflections inside the discs (no bound on a face); the
    # first steps are the 128 nm light's scatterings in the argon, most of them bounded
    c = L.simulate(lar, batch, seed=5, controls=Controls(max_vertices=3, reach_bound=True))
    d = L.simulate(lar, batch, seed=5, controls=Controls(max_vertices=3))
    _same(c, d, "lar, 3 steps")
    print(f"[reach] lar, 3 steps: {c.health.bounded / c.health.vertices:.3f} of the steps bounded", flush=True)
    assert c.health.bounded > 0.5 * c.health.vertices


def test_water_tank_the_same_with_and_without_the_bound(gpu):
    det = water_tank()
    x0, d = muon_track()
    src = L.sources.CherenkovSegment("Water", start=tuple(x0), direction=tuple(d), length=1.0)
    a = L.simulate(det, src, samples=1 << 19, seed=7, controls=ON)
    b = L.simulate(det, src, samples=1 << 19, seed=7)
    _same(a, b, "tank")
    print(f"[reach] tank: {a.health.bounded / a.health.vertices:.4f} of the steps bounded", flush=True)


@pytest.mark.parametrize("name", sorted(S.SCENES))
def test_scenes_the_same_with_and_without_the_bound(gpu, name):
    """Flush, coincident, nested and boolean faces, thin wires: the scenes where a crossing goes on past a face."""
    det = S.detector(name)
    src = L.sources.Point((0.05, 0.03, 0.02), lam=420.0, n_photons=1.0)
    a = L.simulate(det, src, samples=1 << 16, seed=3, controls=ON)
    b = L.simulate(det, src, samples=1 << 16, seed=3)
    _same(a, b, name)


def test_lar_adjoint_the_same_with_and_without_the_bound(gpu, lar):
    src = L.sources.Point((0.5, 0.0, -0.4), lam=128.0, n_photons=1.0)
    p0 = lar.params()
    grads = []
    for ctl in (ON, None):
        tr = table_tensor(p0, "materials/LAr/RAYLEIGH")
        tf = table_tensor(p0, "surfaces/PTFE/REFLECTIVITY")
        light = simulate_diff(lar, src, {"materials/LAr/RAYLEIGH": tr, "surfaces/PTFE/REFLECTIVITY": tf},
                              samples=1 << 16, seed=11, controls=ctl)
        light.arrivals.sum().backward()
        grads.append(np.concatenate([tr.grad.cpu().numpy(), tf.grad.cpu().numpy()]))
    scale = np.abs(grads[1]).max()
    assert scale > 0
    np.testing.assert_allclose(grads[0], grads[1], rtol=1e-6, atol=1e-9 * scale)
