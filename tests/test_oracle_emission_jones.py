"""Pins of the oracle's Cherenkov emitter and of its Jones (polarisation) steps by what the physics fixes, not by the
oracle's own formulas.

Cherenkov (the reference's ``sources/cherenkov.py:33-66`` ``cone``, ``surfaces/polar.py:96-104`` ``cherenkov``):
every photon sits on its cone cos theta = 1/(beta n(lambda)) with n read from the RINDEX table linearly in energy;
the weighted spectrum is Frank-Tamm's dN/dlambda = 2 pi alpha z^2 L (1 - 1/(beta n)^2) / lambda^2 under a dispersive
index; the start is uniform along the track and the time is the charge's passage s/(beta c); the polarisation is
transverse, in the plane of the track and the photon, pointing against the track.

Jones steps (``surfaces/polar.py:57-120``: ``components``, ``recombine``, ``mirror``): a specular reflection and a
refraction recombined from the Fresnel amplitudes satisfy Maxwell's boundary conditions (tangential E and H
continuous, normal D continuous) for any complex incident state; total internal reflection keeps the state's
energy and gives the textbook relative phase of s and p (Born and Wolf 1.5.4); a mirror of normal h leaves E + E'
along h (a perfect conductor's boundary condition) and keeps the state transverse to the mirrored direction.
"""

import math

import numpy as np
import pytest
from scipy import integrate, stats

import lumen_amd as L
from lumen_amd import pack, units
from lumen_amd.detector import water_tank
from oracle import oracle as O

BETA, LENGTH, CHARGE = 0.95, 1.0, 1.0
START, AXIS = np.array([0.3, -0.2, -0.5]), np.array([0.2, -0.3, 0.9]) / np.linalg.norm([0.2, -0.3, 0.9])


@pytest.fixture(scope="module")
def emitted():
    det = water_tank()
    src = L.sources.CherenkovSegment("Water", start=tuple(START), direction=tuple(AXIS), length=LENGTH, beta=BETA,
                                     charge=CHARGE, time=1.5)
    samples = 1 << 17
    ph = O.OracleRun(det).emit(src, samples, samples, seed=3)
    # the RINDEX table as the loop reads it (packed float32), linear in photon energy
    t = det.params()[det.media[[m.name for m in det.media].index("Water")].rindex]
    e = np.asarray(t.energy, np.float32).astype(np.float64)
    nv = np.asarray(t.value, np.float32).astype(np.float64)
    return ph, samples, e, nv


def _n_of(lam, e, nv):
    return np.interp(units.HC / lam, e, nv)


def test_cherenkov_photons_sit_on_their_cone(emitted):
    ph, _, e, nv = emitted
    axis = AXIS.astype(np.float32).astype(np.float64)
    axis /= np.linalg.norm(axis)
    cos = ph["d"] @ axis
    np.testing.assert_allclose(np.linalg.norm(ph["d"], axis=1), 1.0, atol=1e-12)
    np.testing.assert_allclose(cos, 1.0 / (BETA * _n_of(ph["lam"], e, nv)), atol=1e-12)
    # the azimuth about the track is uniform
    t1 = np.cross(axis, [1.0, 0.0, 0.0])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(axis, t1)
    phi = np.arctan2(ph["d"] @ t2, ph["d"] @ t1)
    assert stats.kstest((phi + math.pi) / (2 * math.pi), "uniform").pvalue > 1e-3


def test_cherenkov_spectrum_is_frank_tamms(emitted):
    """The weight-summed photons per wavelength bin against 2 pi alpha z^2 L int (1 - 1/(beta n)^2)/lambda^2 over
    the bin (chi^2 with the weights' own variance), and 1/lambda uniform (the proposal)."""
    ph, samples, e, nv = emitted
    lam_lo, lam_hi = units.HC / e.max(), units.HC / e.min()
    assert ph["lam"].min() >= lam_lo - 1e-9 and ph["lam"].max() <= lam_hi + 1e-9
    inv = (1.0 / ph["lam"] - 1.0 / lam_hi) / (1.0 / lam_lo - 1.0 / lam_hi)
    assert stats.kstest(inv, "uniform").pvalue > 1e-3
    edges = np.linspace(lam_lo, lam_hi, 25)
    per_m = 2.0 * math.pi * units.FINE_STRUCTURE * CHARGE**2 * LENGTH * units.NM_PER_M

    def density(lam):
        n = _n_of(lam, e, nv)
        return per_m * max(1.0 - 1.0 / (BETA * n) ** 2, 0.0) / lam**2

    expect = np.array([integrate.quad(density, a, b, limit=200, points=units.HC / e[(units.HC / e > a) &
                                                                                     (units.HC / e < b)])[0]
                       for a, b in zip(edges[:-1], edges[1:])])
    got = np.histogram(ph["lam"], edges, weights=ph["w0"])[0]
    var = np.histogram(ph["lam"], edges, weights=ph["w0"] ** 2)[0]
    chi2 = float((((got - expect) ** 2) / var).sum())
    assert stats.chi2.sf(chi2, len(expect)) > 1e-3, chi2
    # the total is Frank-Tamm's yield over the table's span
    total = integrate.quad(density, lam_lo, lam_hi, limit=400, points=units.HC / e)[0]
    assert abs(ph["w0"].sum() - total) < 4.0 * math.sqrt((ph["w0"] ** 2).sum())


def test_cherenkov_start_and_time(emitted):
    ph, _, _, _ = emitted
    start = START.astype(np.float32).astype(np.float64)
    axis = AXIS.astype(np.float32).astype(np.float64)
    axis /= np.linalg.norm(axis)
    rel = ph["x"] - start
    s = rel @ axis
    np.testing.assert_allclose(np.linalg.norm(rel - s[:, None] * axis, axis=1), 0.0, atol=1e-6)  # float32 axis
    assert s.min() >= -1e-9 and s.max() <= LENGTH + 1e-9
    assert stats.kstest(s / LENGTH, "uniform").pvalue > 1e-3
    np.testing.assert_allclose(ph["t"], 1.5 + s / (np.float32(BETA) * units.C_VACUUM), rtol=1e-6, atol=1e-9)
    assert np.all(ph["vol"] >= 0)


def test_cherenkov_polarisation(emitted):
    ph, _, _, _ = emitted
    axis = AXIS.astype(np.float32).astype(np.float64)
    axis /= np.linalg.norm(axis)
    er, d = ph["er"], ph["d"]
    np.testing.assert_allclose(np.linalg.norm(er, axis=1), 1.0, atol=1e-12)
    np.testing.assert_allclose((er * d).sum(1), 0.0, atol=1e-12)  # transverse
    np.testing.assert_allclose((er * np.cross(axis, d)).sum(1), 0.0, atol=1e-12)  # in the (axis, d) plane
    assert np.all(er @ axis < 0.0)  # against the track (Geant4's -sin theta component)
    np.testing.assert_allclose(ph["ei"], 0.0, atol=0.0)  # linear


# ------------------------------------------------------------------------------------------------------- Jones
def _random_state(rng, k):
    """A random complex state transverse to k (unnormalised)."""
    a = rng.normal(size=3) + 1j * rng.normal(size=3)
    a = a - (a @ k) * k
    return a


def _tangential(v, n):
    return v - np.outer(v @ n, n).reshape(v.shape) if v.ndim > 1 else v - (v @ n) * n


def _snell(wi, n, n1, n2):
    c = -wi @ n
    eta = n1 / n2
    s2 = eta * eta * (1.0 - c * c)
    return eta * wi + (eta * c - math.sqrt(1.0 - s2)) * n, math.sqrt(1.0 - s2)


@pytest.mark.parametrize("n1,n2", [(1.33, 1.0), (1.0, 1.5), (1.5, 1.33), (1.34, 1.47)])
def test_reflection_and_refraction_satisfy_maxwells_boundary_conditions(n1, n2):
    rng = np.random.default_rng(int(100 * n1 + 10 * n2))
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        wi = rng.normal(size=3)
        wi -= 2.0 * max(wi @ n, 0.0) * n  # against the normal
        wi /= np.linalg.norm(wi)
        c = -wi @ n
        if n1 > n2 and (n1 / n2) ** 2 * (1 - c * c) >= 0.999 or c < 1e-3:
            continue  # total internal reflection (below) or grazing
        E = _random_state(rng, wi)
        rs, rp, ts, tp = O.fresnel(c, n1, n2)
        kr = wi + 2.0 * c * n
        kt, _ = _snell(wi, n, n1, n2)
        er_r, ei_r = O.pol_recombine(wi, n, rs, rp, kr, E.real, E.imag)
        er_t, ei_t = O.pol_recombine(wi, n, ts, tp, kt, E.real, E.imag)
        Er, Et = er_r + 1j * ei_r, er_t + 1j * ei_t
        # the recombined states are normalised: restore each wave's size from the powers of the s and p parts
        s = np.cross(wi, n)
        s /= np.linalg.norm(s)
        es2 = abs(E @ s) ** 2
        ep2 = float(np.vdot(E, E).real) - es2
        Er = Er * math.sqrt(abs(rs) ** 2 * es2 + abs(rp) ** 2 * ep2)
        Et = Et * math.sqrt(abs(ts) ** 2 * es2 + abs(tp) ** 2 * ep2)
        scale = math.sqrt(float(np.vdot(E, E).real))
        assert abs(Er @ kr) < 1e-12 * scale and abs(Et @ kt) < 1e-12 * scale  # transverse waves
        np.testing.assert_allclose(_tangential(E + Er - Et, n), 0.0, atol=1e-12 * scale)  # tangential E
        H = n1 * np.cross(wi, E) + n1 * np.cross(kr, Er) - n2 * np.cross(kt, Et)  # H ~ n k x E (non-magnetic)
        np.testing.assert_allclose(_tangential(H, n), 0.0, atol=1e-12 * scale)  # tangential H
        assert abs(n1**2 * (E + Er) @ n - n2**2 * Et @ n) < 1e-12 * scale  # normal D


def test_total_internal_reflection_phase():
    """Beyond the critical angle the reflected state keeps its energy and the s-p relative phase delta of
    tan(delta / 2) = cos(theta) sqrt(sin^2 theta - n^2) / sin^2 theta (n = n2/n1): for a 45-degree linear state
    the reflected state's circular share |S3| / S0 = 2 |(E_r x E_i) . k| / |E|^2 is |sin delta|, whatever the
    basis's orientation."""
    n1, n2 = 1.5, 1.0
    nrel = n2 / n1
    n = np.array([0.0, 0.0, 1.0])
    for theta in np.radians([42.5, 45.0, 50.0, 55.0, 65.0, 80.0]):
        wi = np.array([math.sin(theta), 0.0, -math.cos(theta)])
        s = np.cross(wi, n)
        s /= np.linalg.norm(s)
        p = np.cross(wi, s)
        E = (s + p) / math.sqrt(2.0)
        rs, rp, _, _ = O.fresnel(math.cos(theta), n1, n2)
        kr = wi + 2.0 * math.cos(theta) * n
        er, ei = O.pol_recombine(wi, n, rs, rp, kr, E, np.zeros(3))
        assert abs(er @ er + ei @ ei - 1.0) < 1e-12 and abs(er @ kr) < 1e-12 and abs(ei @ kr) < 1e-12
        s3 = 2.0 * abs(np.cross(er, ei) @ kr)
        sn2 = math.sin(theta) ** 2
        half = math.atan(math.cos(theta) * math.sqrt(sn2 - nrel**2) / sn2)
        assert abs(s3 - abs(math.sin(2.0 * half))) < 1e-12


def test_normal_incidence_reflects_and_transmits_the_whole_state():
    """At normal incidence no plane of incidence exists: the reflected field is r E and the transmitted t E for any
    state, r = (n1 - n2)/(n1 + n2), t = 2 n1/(n1 + n2) (up to the common normalisation)."""
    rng = np.random.default_rng(4)
    n1, n2 = 1.0, 1.5
    n = np.array([0.0, 0.0, 1.0])
    wi = -n
    for _ in range(20):
        E = _random_state(rng, wi)
        E /= math.sqrt(float(np.vdot(E, E).real))
        rs, rp, ts, tp = O.fresnel(1.0, n1, n2)
        er, ei = O.pol_recombine(wi, n, rs, rp, n, E.real, E.imag)
        r = (n1 - n2) / (n1 + n2)
        np.testing.assert_allclose(er + 1j * ei, np.sign(r) * E, atol=1e-12)
        er, ei = O.pol_recombine(wi, n, ts, tp, wi, E.real, E.imag)
        np.testing.assert_allclose(er + 1j * ei, E, atol=1e-12)


def test_mirror_is_a_perfect_conductor():
    rng = np.random.default_rng(5)
    for _ in range(200):
        wi = rng.normal(size=3)
        wi /= np.linalg.norm(wi)
        wo = rng.normal(size=3)
        wo /= np.linalg.norm(wo)
        h = wo - wi
        h /= np.linalg.norm(h)
        E = _random_state(rng, wi)
        er, ei = O.pol_mirror(h, E.real, E.imag)
        Eo = er + 1j * ei
        np.testing.assert_allclose(_tangential(E + Eo, h), 0.0, atol=1e-12)  # tangential total field vanishes
        assert abs(np.vdot(Eo, Eo) - np.vdot(E, E)) < 1e-12  # no loss
        assert abs(Eo @ wo) < 1e-12  # transverse to the mirrored direction


def test_oracle_cone_toy_closed_form():
    """The oracle's whole loop on the cone toy (a 0.3 m track along +z at beta 0.9 in plain water inside a ring
    counter): each ring's share of the Frank-Tamm yield and its mean arrival time against the geometry's closed
    form (tests/toys.py cone_rings_closed_form)."""
    from toys import cone_rings_closed_form, cone_toy

    rings, beta, n_idx = 90, 0.9, 1.33
    det = cone_toy(rings, n_idx)
    src = L.sources.CherenkovSegment("Water", start=(0.0, 0.0, -0.15), direction=(0.0, 0.0, 1.0), length=0.3,
                                     beta=beta)
    samples = 1 << 16
    tal, h = O.OracleRun(det).trace(src, samples, seed=9)
    o = O.light_of(tal, h, rings, 1, samples)
    total = float(h[14])  # the start weight: Frank-Tamm's yield
    share, tmean = cone_rings_closed_form(rings, -0.15, 0.15, 1.0 / (np.float32(beta) * n_idx), np.float32(beta),
                                          n_idx)
    hit = share > 0
    assert abs(o["arrivals"].sum() / total - 1.0) < 1e-9  # every photon reaches a ring
    err = o["arrivals_error"] / total
    np.testing.assert_allclose(o["arrivals"][~hit], 0.0, atol=0.0)
    pulls = (o["arrivals"][hit] / total - share[hit]) / np.maximum(err[hit], 1e-12)
    assert np.abs(pulls).max() < 5.0, pulls
    big = share > 0.05
    np.testing.assert_allclose(o["mean_arrival_time"][big], tmean[big], rtol=2e-3)


def test_batch_event_error_uses_the_events_photons():
    """A Batch whose event 0 merges two sources (1,000 and 3,000 photons) and whose event 1 is one source (2,000),
    each photon standing for one (weights 1), isotropic at the centre of the two-ring counter: each ring's count in an
    event is binomial (n_e, 1/2), so the moment error must estimate sqrt(n_e / 4) with n_e the event's photons --
    over 24 seeds its mean square matches n_e / 4 and the arrivals' spread."""
    from types import SimpleNamespace

    import torch

    from lumen_amd.api import light_of
    from toys import cone_toy

    det = cone_toy(2)
    srcs = tuple(L.sources.Point((0.0, 0.0, 0.0), lam=400.0, n_photons=float(n)) for n in (1000, 3000, 2000))
    batch = L.sources.Batch(srcs, (1000, 3000, 2000), events=(0, 0, 1))
    eng = SimpleNamespace(S=2, K=1, nt=0, Kl=1, Kt=1)
    orc = O.OracleRun(det)
    errs, arrs = [], []
    for seed in range(24):
        tal, h = orc.trace(batch, None, seed=seed)
        light = light_of(eng, torch.as_tensor(tal), h, None, batch)
        errs.append(light.arrivals_error.numpy())
        arrs.append(light.arrivals.numpy())
    e2, a = np.mean(np.square(errs), 0), np.asarray(arrs)
    n_e = np.array([4000.0, 2000.0])[:, None]
    np.testing.assert_allclose(a.mean(0), n_e / 2.0 * np.ones((1, 2)), rtol=0.02)
    np.testing.assert_allclose(e2, n_e / 4.0 * np.ones((1, 2)), rtol=0.01)  # the estimator's mean, not its spread
    assert np.all(np.abs(a.var(0, ddof=1) / (n_e / 4.0) - 1.0) < 0.8)  # 24 seeds: chi2 with 23 dof
