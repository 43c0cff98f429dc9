"""The reference's navigator edge-case scenes (tests/nav_scenes.py) in the oracle: points locate where the scene puts
them, and photons from inside traverse nested, flush, face-sharing, boolean, assembled, framed and coincident volumes
with no photon lost, abandoned or pushed, the fates balancing; the coincident scene and its variant a few bands
apart give the same light."""

import numpy as np
import pytest

import lumen_amd as L
import nav_scenes as S
from oracle import oracle as O

SRC = L.sources.Point((0.05, 0.03, 0.02), lam=420.0, n_photons=1.0)


@pytest.mark.parametrize("name", sorted(S.SCENES))
def test_scene_traces_cleanly(name):
    det = S.detector(name)
    n = 1 << 13
    tal, h = O.OracleRun(det).trace(SRC, n, seed=3)
    assert h[2] == 0.0 and h[5] == 0.0 and h[9] == 0.0, h  # lost, abandoned, pushed
    fates = h[0] + h[1] + h[3] + h[4] + h[6] + h[7]
    assert abs(fates - h[14]) < 1e-9 * h[14]  # absorbed, escaped, cut, truncated, rouletted: the start weight


def test_nest_locates():
    det = S.detector("nest")
    names = [n for n in det.sensor_names]
    orc = O.OracleRun(det)
    pts = np.array([[0.0, 0.0, 0.45], [-0.05 + 0.1 * np.cos(np.radians(20)), 0.05 + 0.1 * np.sin(np.radians(20)), 0],
                    [-0.7, 0.0, 0.0], [-0.5, 0.05, 0.0], [1.5, 0, 0]])
    v = orc.locate(pts)
    want = ["world/tank_0", None, "world/left_1", "world/right_2", "OUT"]
    for vi, w in zip(v, want):
        if w == "OUT":
            assert vi == det.nav.n
        elif w is not None:
            assert names[det.sensor_of_vol[vi]] == w


def test_coincident_and_near_give_the_same_light():
    a, b = S.detector("coincident"), S.detector("near")
    n = 1 << 13
    la = O.light_of(*O.OracleRun(a).trace(SRC, n, seed=4), a.n_sensors, 1, n)["arrivals"]
    lb = O.light_of(*O.OracleRun(b).trace(SRC, n, seed=4), b.n_sensors, 1, n)["arrivals"]
    np.testing.assert_allclose(la, lb, atol=2e-3)


def test_exit_through_a_shared_face_enters_the_sibling():
    """A beam inside `right` along -x leaves it through the face it shares with `left` (x = -0.6): the crossing goes
    on into `left` (its entry detects a tenth), not into the world between them; `flush`, flush with the tank's +x
    face, is left into the tank and then the world, no photon lost."""
    det = S.detector("nest")
    orc = O.OracleRun(det)
    n = 1 << 10
    names = det.sensor_names
    for fp32 in (False, True):
        src = L.sources.Point((-0.5, 0.01, 0.02), lam=420.0, n_photons=1.0, direction=(-1.0, 0.0, 0.0))
        tal, h = orc.trace(src, n, seed=1, fp32=fp32)
        a = O.light_of(tal, h, det.n_sensors, 1, n)["arrivals"]
        assert abs(a[names.index("world/left_1")] - 0.1) < 1e-6 and a[names.index("world/right_2")] == 0.0
        assert h[2] == 0.0 and h[9] == 0.0


def test_an_entry_through_coincident_faces_is_one_crossing():
    """A beam down onto the coincident scene's top face (z = 0.3, where env, glass and inner are flush) crosses it
    once, from the world into inner (a tenth detected there), then leaves inner into glass at z = 0.2 (a tenth of
    the 0.9 going on): glass's top face is never crossed on its own."""
    for name in ("coincident", "near"):
        det = S.detector(name)
        names = det.sensor_names
        n = 1 << 10
        src = L.sources.Point((0.05, 0.02, 0.5), lam=420.0, n_photons=1.0, direction=(0.0, 0.0, -1.0))
        for fp32 in (False, True):
            tal, h = O.OracleRun(det).trace(src, n, seed=1, fp32=fp32)
            a = O.light_of(tal, h, det.n_sensors, 1, n)["arrivals"]
            assert abs(a[names.index("world/env_0/glass_0/inner_0")] - 0.1) < 1e-6, (name, a)
            assert abs(a[names.index("world/env_0/glass_0")] - 0.09) < 1e-6, (name, a)
