"""The surface types beyond the six of the first build (the reference's ``surfaces/types.py:445-573``): the
photocathode's complex index against the textbook reflectance of an absorbing medium, and every new type on the
sensor sphere (a point at the centre: every photon meets the wall at normal incidence, a mirror or a backscatter sends
it back through the centre to the wall again) against the closed form of its light."""

import math

import numpy as np
import pytest

import lumen_amd as L
from lumen_amd.detector.model import BACKSCATTER, DIFFUSE, SPECULAR, STRAIGHT, SurfaceRow
from oracle import oracle as O
from toys import lambert_closed_form, specular_closed_form, sphere_toy

A, L_ABS, N_IDX = 1.0, 10.0, 1.33
SRC = L.sources.Point((0.0, 0.0, 0.0), lam=420.0, n_photons=1.0)


def _normal_reflectance(n1, nr, nk):
    """R at normal incidence from n1 into nr + i nk: ((n1 - nr)^2 + nk^2) / ((n1 + nr)^2 + nk^2) (textbook)."""
    return ((n1 - nr) ** 2 + nk**2) / ((n1 + nr) ** 2 + nk**2)


@pytest.mark.parametrize("nr,nk", [(2.9, 1.6), (0.2, 3.5), (1.8, 0.0), (1.0, 0.3)])
def test_complex_fresnel_textbook(nr, nk):
    n1 = N_IDX
    rs, rp, _, _ = O.fresnel_complex(1.0, n1, nr, nk)
    rn = _normal_reflectance(n1, nr, nk)
    assert abs(abs(rs) ** 2 - rn) < 1e-12 and abs(abs(rp) ** 2 - rn) < 1e-12
    if nk == 0.0:  # a real index: the dielectric's amplitudes
        for c in (0.2, 0.5, 0.9):
            np.testing.assert_allclose(O.fresnel_complex(c, n1, nr, 0.0)[:2], O.fresnel(c, n1, nr)[:2], atol=1e-12)
    for c in (0.05, 0.3, 0.6, 0.95):  # an absorbing medium reflects s at least as well as p; grazing: both -> 1
        rs, rp, _, _ = O.fresnel_complex(c, n1, nr, nk)
        assert abs(rp) ** 2 <= abs(rs) ** 2 + 1e-12 and abs(rs) <= 1.0 + 1e-12
    rs, rp, _, _ = O.fresnel_complex(1e-6, n1, nr, nk)
    assert abs(rs) ** 2 > 0.999 and abs(rp) ** 2 > 0.999


def _q(det, n=1 << 16, seed=1):
    tal, h = O.OracleRun(det).trace(SRC, n, seed=seed)
    o = O.light_of(tal, h, 1, tal.shape[1] // 1 - 3, n)
    return float(o["arrivals"][0]), float(o["arrivals_error"][0]), h


def test_photocathode_sphere_closed_form():
    """The cathode's reflectance at normal incidence is the mirror's R: Q/N = eff (1 - R) e^(-a/L) / (1 - R e^(-2a/L))."""
    nr, nk, eff = 2.9, 1.6, 0.8
    det = sphere_toy(A, "PHOTOCATHODE", R=0.0, L_abs=L_ABS, n=N_IDX, eff=eff, extra={"nr": nr, "nk": nk})
    q, err, h = _q(det)
    rn = _normal_reflectance(N_IDX, nr, nk)
    want = eff * specular_closed_form(A, rn, L_ABS)
    assert abs(q - want) < 1e-6 * want + 5 * err, (q, want, err)
    assert abs(h[14] - sum(h[:9])) < 1e-9 * h[14]
    with pytest.raises(ValueError, match="no derivative"):
        L.pack.pack(det, grad_keys=["surfaces/Wall/nr"])


def test_counting_sphere_closed_form():
    det = sphere_toy(A, "COUNTING", R=0.7, L_abs=L_ABS, n=N_IDX, eff=0.6)
    q, err, _ = _q(det)
    assert abs(q - 0.6 * math.exp(-A / L_ABS)) < 1e-6  # every photon absorbed at its first crossing: no variance


@pytest.mark.parametrize("surface,cols", [("SPECULAR_METAL", {}), ("UNIFIED_METAL_NORMAL",
                                                                    {"p_spike": 0.3, "p_lobe": 0.3, "p_back": 0.4})])
def test_transmitting_mirror_sphere_closed_form(surface, cols):
    """R back through the centre (spike, lobe about the normal, backscatter alike), T straight into the opaque shell
    (absorbed, not detected): Q/N = eff (1 - R - T) e^(-a/L) / (1 - R e^(-2a/L))."""
    R, T, eff = 0.5, 0.2, 0.9
    det = sphere_toy(A, surface, R=R, L_abs=L_ABS, n=N_IDX, eff=eff, extra=dict(cols, T=T), transmitting=True,
                     far_opaque=True)
    q, err, h = _q(det)
    want = eff * (1 - R - T) * math.exp(-A / L_ABS) / (1 - R * math.exp(-2 * A / L_ABS))
    assert abs(q - want) < 1e-6 * want + 5 * err, (q, want, err)


def test_painted_t_sphere_closed_form():
    """Lambertian R (f_spec 0), T straight into the opaque shell: the Lambertian sphere's closed form with the
    absorbed share 1 - R - T detected."""
    R, T, eff = 0.6, 0.15, 1.0
    det = sphere_toy(A, "PAINTED_T", R=R, L_abs=L_ABS, n=N_IDX, eff=eff, extra={"T": T}, far_opaque=True)
    q, err, _ = _q(det, n=1 << 18)
    b = 2 * A / L_ABS
    want = (1 - R - T) * math.exp(-A / L_ABS) / (1 - R * 2 * (1 - (1 + b) * math.exp(-b)) / b**2)
    assert abs(q - want) < 5 * err, (q, want, err)
    assert abs(lambert_closed_form(A, R, L_ABS) * (1 - R - T) / (1 - R) - want) < 1e-12


def test_support_rules_of_fixed_constants():
    """The defensive floor covers only what the row's fixed constants leave possible."""
    def support(row):
        return {k for k in range(6) if (row.flags() >> 8) >> k & 1}

    assert support(SurfaceRow("u", "UNIFIED_METAL_NORMAL", {"R": "k", "p_spike": 0.0, "p_lobe": 0.0, "p_back": "b"})) \
        == {BACKSCATTER, DIFFUSE}
    assert support(SurfaceRow("u", "UNIFIED_METAL_NORMAL", {"R": "k", "p_spike": 0.5, "p_lobe": 0.5, "p_back": 0.0})) \
        == {SPECULAR}
    assert support(SurfaceRow("p", "PAINTED_T", {"R": "k", "T": 0.0, "p_spike": 1.0})) == {SPECULAR}
    assert support(SurfaceRow("p", "PAINTED_T", {"R": "k", "T": "t", "p_spike": 0.0})) == {STRAIGHT, DIFFUSE}
    assert support(SurfaceRow("s", "SPECULAR_METAL", {"R": "k", "T": "t"}, transmitting=True)) == {SPECULAR, STRAIGHT}
    assert support(SurfaceRow("c", "COUNTING", {"eff": "e"})) == set()
    assert support(SurfaceRow("pc", "PHOTOCATHODE", {"nr": "a", "nk": "b", "eff": "e"})) == {SPECULAR}
    with pytest.raises(ValueError):
        SurfaceRow("x", "PAINTED", {}, transmitting=True)


@pytest.mark.parametrize("surface,extra", [("PHOTOCATHODE", {"nr": 2.9, "nk": 1.6}), ("SPECULAR_METAL", {})])
def test_per_normal_detection_is_eff_at_normal_incidence(surface, extra):
    """Read for a QE applied after tracking, a crossing's detected share is over A_n (the absorbed share at normal
    incidence): at the sphere's normal incidence every crossing detects eff of what reaches it,
    Q/N = eff e^(-a/L) / (1 - R_n e^(-2a/L))."""
    eff, R = 0.7, 0.5
    det = sphere_toy(A, surface, R=R, L_abs=L_ABS, n=N_IDX, eff=eff, extra=extra, per_normal=True)
    q, err, _ = _q(det)
    rn = _normal_reflectance(N_IDX, extra["nr"], extra["nk"]) if extra else R
    want = eff * math.exp(-A / L_ABS) / (1 - rn * math.exp(-2 * A / L_ABS))
    assert abs(q - want) < 1e-6 * want + 5 * err, (q, want, err)
