"""The reference's navigator edge-case scenes on the device (tests/nav_scenes.py: nest, bool, wires, coincident,
near): the float32 navigator's next crossing against the FP64 referee on random rays and on rays aimed at the
scenes' shared, flush and coincident faces (distance within 1e-5 relative plus a few float32 ulps of the
coordinates), and the traced light and fates against the oracle's photons."""

import numpy as np
import pytest

import lumen_amd as L
import nav_scenes as S
from lumen_amd.api import engine
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# points on the faces the scenes share: (plane axis, value, extent of the other two coordinates)
FACES = {
    "nest": [(0, 0.5 - 0.1 + 0.1, (-0.3, 0.3))] + [(0, -0.6, (-0.1, 0.1))],
    "bool": [(0, -0.4 + 0.2, (-0.2, 0.2)), (0, 0.4 + 0.15, (-0.05, 0.05))],
    "wires": [(0, 0.01, (-0.3, 0.3)), (0, -0.01, (-0.3, 0.3))],
    "coincident": [(0, 0.3, (-0.3, 0.3)), (1, -0.3, (-0.3, 0.3)), (2, 0.3, (-0.3, 0.3)), (0, 1.0, (-0.2, 0.2)),
                   (0, 0.1, (-0.1, 0.1))],
}
FACES["near"] = FACES["coincident"]


def _rays(name, rng, n):
    x = rng.uniform(-0.99, 0.99, (n, 3))
    d = rng.normal(size=(n, 3))
    faces = FACES[name]
    m = n // 2
    tgt = np.zeros((m, 3))
    for i in range(m):  # aimed at a point of a shared face
        ax, v, (lo, hi) = faces[i % len(faces)]
        tgt[i] = rng.uniform(lo, hi, 3)
        tgt[i, ax] = v
    d[:m] = tgt - x[:m]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return x.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("name", sorted(S.SCENES))
def test_scene_distances_against_fp64(gpu, name):
    print(f"[nav scenes] {name}", flush=True)
    det = S.detector(name)
    x, d = _rays(name, np.random.default_rng(29), 8192)
    orc = O.OracleRun(det)
    vol = orc.locate(x)
    t64, va64, _, _ = orc.query(x, d, vol)
    t32, va32, _, _, loc = engine(det).query(x, d, vol)
    assert np.mean(loc == vol) > 0.999
    same = va32 == va64
    err = np.abs(t32[same] - t64[same])
    good = err <= 1e-5 * t64[same] + 4e-7
    print(f"[nav scenes] {name}: same volume {same.mean():.5f}, within bound {good.mean():.5f}, max {err.max():.2e}")
    assert same.mean() >= 0.998
    assert good.mean() >= 0.999 and err.max() < 1e-4


@pytest.mark.parametrize("name", sorted(S.SCENES))
def test_scene_traces_against_oracle(gpu, name):
    det = S.detector(name)
    src = L.sources.Point((0.05, 0.03, 0.02), lam=420.0, n_photons=1.0)
    n = 1 << 16
    g = L.simulate(det, src, samples=n, seed=3)
    tal, h = O.OracleRun(det).trace(src, n, seed=3)
    o = O.light_of(tal, h, det.n_sensors, 1, n)
    np.testing.assert_allclose(g.arrivals.cpu().numpy(), o["arrivals"], rtol=5e-3, atol=2e-5)
    hh = g.health
    assert hh.lost == 0.0 and hh.abandoned == 0.0 and hh.pushed <= 1e-4
    assert abs(hh.balance()) < 1e-5 * hh.start_weight
