"""Parity of the HIP engine with the oracle (FP64 CPU restatement), the reference's published numbers and the
closed forms, through the C ABI (``lh_trace``, ``lh_query``)."""

import json
import math
import os

import numpy as np
import pytest
from scipy import stats

import lumen_amd as L
from lumen_amd import units
from lumen_amd.api import engine
from lumen_amd.detector import example_tank, muon_track, water_tank
from oracle import oracle as O
from toys import lambert_closed_form, specular_closed_form, sphere_toy

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
SRC1 = L.sources.Point((0.0, 0.0, 0.1), lam=400.0, n_photons=2e5)


@pytest.fixture(scope="module")
def tank1():
    return example_tank()


@pytest.fixture(scope="module")
def tank2():
    return water_tank()


def _cherenkov():
    x0, d = muon_track()
    return L.sources.CherenkovSegment("Water", start=tuple(x0), direction=tuple(d), length=1.0)


def _np(light):
    return {k: getattr(light, k).detach().cpu().numpy() for k in ("arrivals", "arrivals_error", "mean_arrival_time",
                                                                    "arrival2", "charge")}


def test_example_tank_same_photons_as_oracle(gpu, tank1):
    """One seed: the engine (float32) and the oracle (FP64) trace the same photons (one Philox contract), so each
    PMT's arrivals agree far inside their Monte Carlo error."""
    n = 1 << 18
    g = _np(L.simulate(tank1, SRC1, samples=n, seed=0))
    tal, h = O.OracleRun(tank1).trace(SRC1, n, seed=0)
    o = O.light_of(tal, h, 8, g["charge"].shape[1], n)
    np.testing.assert_allclose(g["arrivals"], o["arrivals"], rtol=2e-4)
    np.testing.assert_allclose(g["arrival2"], o["arrival2"], rtol=2e-3)
    np.testing.assert_allclose(g["mean_arrival_time"], o["mean_arrival_time"], rtol=1e-4)


def test_example_tank_published_arrivals(gpu, tank1):
    tut = json.load(open(os.path.join(GOLDEN, "tutorial_example_tank.json")))
    g = _np(L.simulate(tank1, SRC1, samples=1 << 18, seed=0))
    pub, perr = np.array(tut["arrivals"]), np.array(tut["arrivals_error"])
    pulls = (g["arrivals"] - pub) / np.hypot(g["arrivals_error"], perr)
    assert stats.chi2.sf(np.sum(pulls**2), 8) > 0.01
    np.testing.assert_allclose(g["mean_arrival_time"], tut["mean_arrival_time_ns"], atol=0.02)
    light = L.simulate(tank1, SRC1, samples=1 << 18, seed=0)
    pe = L.readout.pe(light, tank1.params(), tank1).cpu().numpy()
    np.testing.assert_allclose(pe / g["arrivals"], 0.245, rtol=1e-6)


def _pulls(a, ea, b, eb):
    ok = (ea > 0) & (eb > 0)
    return (a[ok] - b[ok]) / np.hypot(ea[ok], eb[ok])


def test_water_tank_pulls_against_the_oracle(gpu, tank2):
    """1,000 PMTs, independent seeds on the two sides, four pairs: the pulls have mean 0 +- 0.05 and width
    1 +- 0.05, and chi2/ndf p > 0.01.  Both sides trace the same number of photons: each PMT's light is skewed
    (a few high-weight photons), so unequal sample sizes bias the pulls' mean by itself (0.26 for 2^20 against
    2^18, measured oracle against oracle)."""
    src = _cherenkov()
    orc = O.OracleRun(tank2)
    pulls = []
    n = 1 << 20
    for k in range(4):
        light = L.simulate(tank2, src, samples=n, seed=1000 + k)
        assert light.health.moment_spills == 0  # no photon deposits at more sensors than its moment slots hold
        g = _np(light)
        tal, h = orc.trace(src, n, seed=2000 + k, bvh=True)  # the BVH walk: brute force's result, 30x faster
        o = O.light_of(tal, h, 1000, 1, n)
        pulls.append(_pulls(g["arrivals"], g["arrivals_error"], o["arrivals"], o["arrivals_error"]))
    p = np.concatenate(pulls)
    assert len(p) > 3900
    assert abs(p.mean()) < 0.05, p.mean()
    assert abs(p.std() - 1.0) < 0.05, p.std()
    assert stats.chi2.sf(np.sum(p**2), len(p)) > 0.01


def test_water_tank_same_photons_as_oracle(gpu, tank2):
    """One seed, the same photons: totals agree to a small fraction of the Monte Carlo error (the facet path's
    float32 rejection can send a rare photon another way)."""
    src = _cherenkov()
    n = 1 << 17
    g = _np(L.simulate(tank2, src, samples=n, seed=7))
    tal, h = O.OracleRun(tank2).trace(src, n, seed=7)
    o = O.light_of(tal, h, 1000, 1, n)
    tot_err = math.sqrt(o["arrival2"].sum())
    assert abs(g["arrivals"].sum() - o["arrivals"].sum()) < 0.05 * tot_err
    assert np.corrcoef(g["arrivals"], o["arrivals"])[0, 1] > 0.999


@pytest.mark.parametrize("which", ["example_tank", "water_tank"])
def test_navigator_distances_against_fp64(gpu, which, tank1, tank2):
    """Fixed rays from random interior points: the device navigator's distance to the next boundary agrees with
    the FP64 referee within 1e-5 relative; at most 0.2 % of rays (grazing ones) meet another volume
    (tests/unit/test_navigator.py:43-73)."""
    det = tank1 if which == "example_tank" else tank2
    rng = np.random.default_rng(11)
    n = 4096
    if which == "example_tank":
        x = rng.uniform(-0.45, 0.45, (n, 3))
    else:
        r = 1.9 * np.sqrt(rng.uniform(0, 1, n))
        ph = rng.uniform(0, 2 * np.pi, n)
        x = np.stack([r * np.cos(ph), r * np.sin(ph), rng.uniform(-1.4, 1.4, n)], 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x32 = x.astype(np.float32).astype(np.float64)
    d32 = d.astype(np.float32).astype(np.float64)
    orc = O.OracleRun(det)
    vol = orc.locate(x32)
    t64, va64, n64, _ = orc.query(x32, d32, vol)
    t32, va32, n32, _, loc = engine(det).query(x32, d32, vol)
    assert np.sum(loc != vol) <= 2
    same = va32 == va64
    assert same.mean() >= 0.998
    rel = np.abs(t32[same] - t64[same]) / t64[same]
    assert rel.max() < 1e-5, rel.max()
    cosn = np.sum(n32[same] * n64[same], 1)
    assert cosn.min() > 1 - 1e-5


def test_specular_sphere_closed_form(gpu):
    a, R, La = 1.0, 0.5, 10.0
    det = sphere_toy(a=a, R=R, L_abs=La)
    g = L.simulate(det, L.sources.Point((0, 0, 0), lam=400.0, n_photons=1.0), samples=1 << 16, seed=2,
                   controls=L.Controls(roulette=1e-12))
    q = float(g.arrivals[0])
    assert abs(q - specular_closed_form(a, R, La)) < 3e-5 * specular_closed_form(a, R, La)
    assert abs(g.health.balance()) < 1e-5


def test_lambert_sphere_and_rayleigh_enclosure(gpu):
    n = 1 << 20
    det = sphere_toy(a=1.0, surface="PAINTED", R=0.6, L_abs=4.0)
    g = L.simulate(det, L.sources.Point((0, 0, 0), lam=400.0, n_photons=1.0), samples=n, seed=3)
    assert abs(float(g.arrivals[0]) - lambert_closed_form(1.0, 0.6, 4.0)) < 4 * float(g.arrivals_error[0])
    det = sphere_toy(a=1.0, R=0.0, L_abs=None, L_ray=0.3)
    g = L.simulate(det, L.sources.Point((0, 0, 0), lam=400.0, n_photons=1.0), samples=n, seed=4)
    assert abs(float(g.arrivals[0]) - 1.0) < max(4 * float(g.arrivals_error[0]), 1e-5)
    assert g.health.vertices > 3 * n


def test_group_velocity_timing(gpu):
    det = sphere_toy(a=1.5, R=0.0, L_abs=None, n=1.40)
    g = L.simulate(det, L.sources.Point((0, 0, 0), lam=400.0, n_photons=1.0), samples=4096, seed=5)
    assert abs(float(g.mean_arrival_time[0]) / (1.5 / (L.units.C_VACUUM / 1.40)) - 1) < 1e-6


def test_photons_do_not_depend_on_the_split(gpu, tank2):
    """Photon i is the same in one launch or in two (first_id), and two runs of one seed agree (FP64 tallies)."""
    src = _cherenkov()
    n = 1 << 18
    a = L.simulate(tank2, src, samples=n, seed=21)
    b = L.simulate(tank2, src, samples=n, seed=21)
    np.testing.assert_allclose(a.arrivals.cpu().numpy(), b.arrivals.cpu().numpy(), rtol=1e-12, atol=1e-15)
    # halves: the same Monte Carlo photons (ids 0..n/2-1 and n/2..n-1), each half's weight over the whole run
    half = n // 2
    rows = [L.sources.CherenkovSegment(**{k: getattr(src, k) for k in ("material", "start", "direction", "length")})]
    h1 = L.simulate(tank2, rows[0], samples=half, seed=21, first_id=0)
    h2 = L.simulate(tank2, rows[0], samples=half, seed=21, first_id=half)
    np.testing.assert_allclose(0.5 * (h1.arrivals + h2.arrivals).cpu().numpy(), a.arrivals.cpu().numpy(),
                               rtol=1e-9, atol=1e-12)


def test_health_balances(gpu, tank2):
    g = L.simulate(tank2, _cherenkov(), samples=1 << 20, seed=3)
    assert abs(g.health.balance()) < 1e-5 * g.health.start_weight
    assert g.health.lost == 0 and g.health.abandoned == 0 and g.health.truncated == 0


def test_reference_values_hold_the_draws(gpu, tank1):
    """reference=params is the plain run; with the reference held the light is smooth in a value (common random
    numbers): the central difference of the total over ABSLENGTH's scale is close to both one-sided ones."""
    p0 = tank1.params()
    n = 1 << 18
    plain = L.simulate(tank1, SRC1, samples=n, seed=0)
    same = L.simulate(tank1, SRC1, p0, samples=n, seed=0, reference=p0)
    np.testing.assert_allclose(plain.arrivals.cpu().numpy(), same.arrivals.cpu().numpy(), rtol=1e-6)
    tot = lambda s: float(L.simulate(tank1, SRC1, p0.scale("materials/Water/ABSLENGTH", s), samples=n, seed=0,  # noqa
                                     reference=p0).arrivals.sum())
    f0, fp, fm = tot(1.0), tot(1.01), tot(0.99)
    d_c, d_p, d_m = (fp - fm) / 0.02, (fp - f0) / 0.01, (f0 - fm) / 0.01
    assert d_c > 0 and abs(d_p - d_c) < 0.05 * d_c and abs(d_m - d_c) < 0.05 * d_c


def test_time_histogram_against_oracle(gpu, tank2):
    """The timed 1,000-PMT tank: the engine's time histogram and overflow cells against the oracle's (the same
    photons), and the readout's density integrating to the pe over a gate that holds them."""
    det = tank2.timed(0.0, 0.5, 80)
    src = _cherenkov()
    n = 1 << 16
    g = L.simulate(det, src, samples=n, seed=4)
    tal, h = O.OracleRun(det).trace(src, n, seed=4, bvh=True)
    o = O.light_of(tal, h, det.n_sensors, 1, n)
    Hg, Ho = g.time.cpu().numpy(), o["time"]
    assert Hg.sum() > 0
    np.testing.assert_allclose(Hg.sum(), Ho.sum(), rtol=2e-3)
    np.testing.assert_allclose(g.late.cpu().numpy().sum(0), o["late"].sum(0), rtol=2e-3, atol=1e-9)
    tot = Hg.sum(-1) + g.late.cpu().numpy()[:, 0] + g.early.cpu().numpy()[:, 0]
    np.testing.assert_allclose(tot, g.arrivals.cpu().numpy(), rtol=1e-6, atol=1e-12)
    p = det.params()
    dens = L.readout.pe_density(g, p, gate=(-10.0, 200.0), bin_width=1.0, tts_sigma=1.3, det=det)
    pe = L.readout.pe(g, p, det)
    np.testing.assert_allclose(dens.sum(-1).cpu().numpy(), pe.cpu().numpy(), rtol=1e-6, atol=1e-12)


def test_time_histogram_moments_on_device(gpu):
    """Every photon of the sensor sphere arrives at T = a / v_g: the node masses' first three moments are T's."""
    a, n_idx = 1.5, 1.40
    T = a / (units.C_VACUUM / n_idx)
    det = sphere_toy(a=a, R=0.0, L_abs=None, n=n_idx).timed(0.0, 0.37, 40)
    g = L.simulate(det, L.sources.Point((0, 0, 0), lam=400.0, n_photons=1.0), samples=1 << 14, seed=2)
    H, tn = g.time.cpu().numpy()[0], 0.37 * np.arange(40)
    s = float(g.arrivals[0])
    np.testing.assert_allclose(H.sum(), s, rtol=1e-6)
    for k in (1, 2, 3):
        np.testing.assert_allclose((H * tn**k).sum() / s, T**k, rtol=2e-5)  # float32 times


CE_KNOTS = np.array([0.0, 0.5, 1.2, 2.0, 2.6, np.pi])  # rad
CE_VALUES = np.array([1.0, 0.9, 0.7, 0.5, 0.3, 0.2])


def test_theta_axis_against_oracle_and_readout(gpu):
    """The charge histogram's theta axis (the cathode position angle, hats on the CE knots): a reflecting sensor
    sphere with absorption traced by the engine and the oracle on the same photons (cells within 1e-4 of the total);
    readout.pe multiplies CE(theta) in (sum_j cell_j CE_j); the time histogram is weighted by CE."""
    from lumen_amd.params import Table
    from lumen_amd.readout import pe, time_masses

    det = sphere_toy(a=1.0, R=0.3, L_abs=2.0, ce=Table(CE_KNOTS, CE_VALUES)).timed(0.0, 0.5, 60)
    src = L.sources.Point((0.2, -0.1, 0.3), lam=400.0, n_photons=1.0)
    n = 1 << 18
    g = L.simulate(det, src, samples=n, seed=4)
    assert g.charge_axes == (1, len(CE_KNOTS))
    tal, h = O.OracleRun(det).trace(src, n, seed=4)
    o = O.light_of(tal, h, 1, len(CE_KNOTS), n)
    gc = g.charge.cpu().numpy().reshape(-1)
    assert np.all(np.abs(gc - o["charge"][0]) < 1e-4 * o["arrivals"][0]), (gc, o["charge"][0])
    assert abs(gc.sum() - float(g.arrivals[0])) < 1e-6 * float(g.arrivals[0])
    p = pe(g, det.params(), det)
    want = float(gc @ CE_VALUES.astype(np.float32))
    assert float(p[0]) == pytest.approx(want, rel=1e-6)
    H, late, early = time_masses(g, det.params(), det)
    assert float(H.sum() + late[..., 0].sum() + early[..., 0].sum()) == pytest.approx(want, rel=1e-5)
    assert float(H.sum()) > 0.9 * want  # the grid holds most of the light


def test_theta_axis_adjoint(gpu):
    """d pe / d(absorption scale) with a CE table: the adjoint's cotangent of each deposit reads its theta cells;
    against common-random-number central differences."""
    from lumen_amd.autograd import simulate_diff, table_tensor
    from lumen_amd.params import Table
    from lumen_amd.readout import pe

    det = sphere_toy(a=1.0, R=0.3, L_abs=2.0, ce=Table(CE_KNOTS, CE_VALUES))
    src = L.sources.Point((0.2, -0.1, 0.3), lam=400.0, n_photons=1.0)
    p0 = det.params()
    key = "materials/Water/ABSLENGTH"
    n = 1 << 18
    t = table_tensor(p0, key)
    light = simulate_diff(det, src, {key: t}, samples=n, seed=6)
    pe(light, p0, det).sum().backward()
    base = np.asarray(p0[key].value, np.float64)
    g = float((t.grad.cpu().numpy() * base).sum())

    def tot(scale):
        p = p0.set(key, L.params.Table(p0[key].energy, base * scale))
        return float(pe(L.simulate(det, src, p, samples=n, seed=6, reference=p0), p, det).sum())

    h = 1e-3
    fd = (tot(1 + h) - tot(1 - h)) / (2 * h)
    assert abs(g - fd) <= 0.01 * abs(fd) + 1e-5 and abs(fd) > 1e-2, (g, fd)


def test_expected_waveform_of_a_simulated_light(gpu):
    """The expected waveform of a traced light (the sensor sphere, times on a grid): its integral over the gate is
    gain m1 pe times the pulse's area sqrt(pi / 2)(sigma_rise + sigma_fall), and it peaks t_peak after the light."""
    from lumen_amd.readout import expected_waveform, pe

    a, n_idx = 1.0, 1.33
    det = sphere_toy(a=a, R=0.0, L_abs=None, n=n_idx).timed(0.0, 0.25, 200)
    light = L.simulate(det, L.sources.Point((0, 0, 0), lam=400.0, n_photons=1000.0), samples=1 << 16, seed=3)
    pulse = dict(t_peak=8.0, sigma_rise=1.5, sigma_fall=4.0)
    w = expected_waveform(light, det.params(), (0.0, 60.0), 0.5, 1.0, pulse, gain=2.0, charge_moments=(1.2, 1.6),
                          baseline=10.0, det=det)
    area = float((w.mean[0] - 10.0).sum()) * w.dt
    want = 2.0 * 1.2 * float(pe(light, det.params(), det)[0]) * math.sqrt(math.pi / 2) * (1.5 + 4.0)
    assert area == pytest.approx(want, rel=2e-3)
    t_peak = w.t0 + w.dt * int(torch_argmax(w.mean[0]))
    assert abs(t_peak - (a * n_idx / units.C_VACUUM + 8.0)) < 1.0


def torch_argmax(x):
    return int(x.argmax())


def test_navigator_on_cones_ellipsoids_and_trds(gpu, tmp_path):
    """The navigator on quadrics with negative eigenvalues (cones), phi wedges (half-plane leaves), a polycone's
    union of frusta, an ellipsoid and a trd (read from GDML): distances to the next boundary from random points
    along random rays against the FP64 referee within 1e-5 relative."""
    from lumen_amd.detector.gdml import Recipe, load
    from toys import GDML_SOLIDS

    f = tmp_path / "solids.gdml"
    f.write_text(GDML_SOLIDS)
    det = load(str(f), Recipe("solids", "World"))
    rng = np.random.default_rng(17)
    n = 8192
    centres = np.array([(-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 1, 0), (0, 0, 1.2)], float)
    x = centres[rng.integers(0, 5, n)] + rng.uniform(-0.55, 0.55, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x32 = x.astype(np.float32).astype(np.float64)
    d32 = d.astype(np.float32).astype(np.float64)
    orc = O.OracleRun(det)
    vol = orc.locate(x32)
    assert len(np.unique(vol)) == 6  # every solid and the world
    t64, va64, n64, _ = orc.query(x32, d32, vol)
    t32, va32, n32, _, loc = engine(det).query(x32, d32, vol)
    assert np.mean(loc == vol) > 0.999
    same = va32 == va64
    assert same.mean() >= 0.995
    # 1e-5 relative, with the float32 coordinates' own absolute floor (a few ulps of ~1 m) for short distances
    err = np.abs(t32[same] - t64[same])
    good = err <= 1e-5 * t64[same] + 2e-6
    assert good.mean() >= 0.999 and err.max() < 1e-4, (err.max(), good.mean())


def test_navigator_and_tallies_on_frame_groups(gpu, tmp_path):
    """Plates sharing a rotation walked in their group's frame on the device: distances along random rays against
    the FP64 referee, and the tallies of the same photons against the oracle's."""
    from test_oracle import _framed_detector

    det = _framed_detector(tmp_path)
    assert det.nav.stats["framed"] == 1
    rng = np.random.default_rng(23)
    n = 8192
    x = rng.uniform(-1.6, 1.6, (n, 3)) * np.array([1.0, 1.0, 0.2])  # the plates' layer, z in [-0.32, 0.32]
    grid = np.array([(a, b, 0.0) for a in (-1.2, -0.4, 0.4, 1.2) for b in (-1.2, -0.4, 0.4, 1.2)])
    x[:1024] = grid[rng.integers(0, 16, 1024)] + rng.uniform(-0.005, 0.005, (1024, 3))  # starts inside a plate
    d = rng.normal(size=(n, 3))
    aim = grid[rng.integers(0, 16, 3072)] + rng.uniform(-0.18, 0.18, (3072, 1)) * np.array([0.7071, 0.7071, 0.0])
    d[1024:4096] = aim - x[1024:4096]  # aimed along a plate's long axis
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x32, d32 = x.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)
    orc = O.OracleRun(det)
    vol = orc.locate(x32)
    t64, va64, _, _ = orc.query(x32, d32, vol)
    t32, va32, _, _, loc = engine(det).query(x32, d32, vol)
    assert np.mean(loc == vol) > 0.999 and np.all((vol[:1024] >= 1) & (vol[:1024] <= 16))
    hit_plate = (va64 >= 1) & (va64 <= 16)
    assert hit_plate.sum() > 1000
    same = va32 == va64
    assert same.mean() >= 0.999
    err = np.abs(t32[same] - t64[same])
    assert np.mean(err <= 1e-5 * t64[same] + 2e-6) >= 0.999 and err.max() < 1e-4, err.max()
    src = L.sources.Point((0.05, 0.1, 0.6), lam=420.0, n_photons=1.0)
    m = 1 << 16
    g = L.simulate(det, src, samples=m, seed=6)
    tal, h = orc.trace(src, m, seed=6, bvh=True)
    o = O.light_of(tal, h, det.n_sensors, 1, m)
    np.testing.assert_allclose(g.arrivals.cpu().numpy(), o["arrivals"], rtol=5e-3, atol=1e-9)


def test_enqueued_runs_keep_their_own_inputs(gpu, tank2):
    """``simulate`` returns before its kernel ends: calls made back to back (other values, other sources, other
    sizes) leave each run's inputs and results its own, equal to the same run made alone; the kernel times of
    every launch stay readable afterwards."""
    import torch

    src, p0 = _cherenkov(), tank2.params()
    p1 = p0.scale("materials/Water/ABSLENGTH", 0.5)
    n = 1 << 20
    runs = [L.simulate(tank2, src, p0, samples=n, seed=11),
            L.simulate(tank2, src, p1, samples=n // 2, seed=12),
            L.simulate(tank2, SRC1, p0, samples=n, seed=13),
            L.simulate(tank2, src, p1, samples=n, seed=11)]
    torch.cuda.synchronize()
    alone = [L.simulate(tank2, src, p0, samples=n, seed=11), L.simulate(tank2, src, p1, samples=n // 2, seed=12),
             L.simulate(tank2, SRC1, p0, samples=n, seed=13), L.simulate(tank2, src, p1, samples=n, seed=11)]
    for a, b in zip(runs, alone):
        np.testing.assert_allclose(a.arrivals.cpu().numpy(), b.arrivals.cpu().numpy(), rtol=1e-9, atol=1e-12)
        # the health cells are float32 per-lane sums added atomically in any order: equal to float32 rounding
        np.testing.assert_allclose(np.asarray(a.health), np.asarray(b.health), rtol=1e-6, atol=1e-6)
        assert a.kernel_ms > 0.0
    assert float(runs[0].arrivals.sum()) > float(runs[3].arrivals.sum())  # the shorter absorption length


def test_runs_on_two_streams_keep_their_own_inputs(gpu, tank2):
    """``simulate(stream=...)`` on two side streams (default priority) back to back, with no wait between: the
    launches share the model's staged sources and work queue, so each waits for the previous one's end; each run
    equals the same run made alone on the current stream."""
    import torch

    src, p0 = _cherenkov(), tank2.params()
    p1 = p0.scale("materials/Water/ABSLENGTH", 0.5)
    n = 1 << 19
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    a = L.simulate(tank2, src, p0, samples=n, seed=31, stream=s1)
    b = L.simulate(tank2, SRC1, p1, samples=n // 2, seed=32, stream=s2)
    c = L.simulate(tank2, src, p1, samples=n, seed=33, stream=s1)
    got = [x.arrivals.cpu().numpy() for x in (a, b, c)]  # read on the current stream, which waits for each
    torch.cuda.synchronize()
    alone = [L.simulate(tank2, src, p0, samples=n, seed=31), L.simulate(tank2, SRC1, p1, samples=n // 2, seed=32),
             L.simulate(tank2, src, p1, samples=n, seed=33)]
    for x, y, z in zip(got, alone, (a, b, c)):
        np.testing.assert_allclose(x, y.arrivals.cpu().numpy(), rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(np.asarray(z.health), np.asarray(y.health), rtol=1e-6, atol=1e-6)


def test_navigator_on_tori(gpu, tmp_path):
    """Torus leaves (a ring, a hollow rotated ring cut in phi, a fat ring) walked by the torus kernels: distances
    along random rays, some aimed through the rings' centres, against the FP64 referee within 1e-5 relative."""
    from test_oracle import _tori

    det = _tori(tmp_path)
    rng = np.random.default_rng(29)
    n = 8192
    centres = np.array([(-0.9, -0.9, 0.0), (0.9, 0.9, 0.0), (0.9, -0.9, 0.8)])
    pick = rng.integers(0, 3, n)
    x = centres[pick] + rng.uniform(-0.75, 0.75, (n, 3))
    d = rng.normal(size=(n, 3))
    d[:2048] = centres[pick[:2048]] + rng.normal(0.0, 0.05, (2048, 3)) - x[:2048]  # through a ring's hole
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x32, d32 = x.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)
    orc = O.OracleRun(det)
    vol = orc.locate(x32)
    assert len(np.unique(vol)) == 4  # the three tori and the world
    t64, va64, n64, _ = orc.query(x32, d32, vol)
    t32, va32, _, _, loc = engine(det).query(x32, d32, vol)
    assert np.mean(loc == vol) > 0.999
    # a torus root met at a grazing angle is conditioned to sqrt(eps) of the tube (the quartic's double root), so the
    # 1e-5 bound is asserted on every ray meeting its face at |cos| >= 0.05, and the grazing rest is held to the
    # float32 root's own bound, 4 sqrt(eps32 t r) for a tube of radius r <= 0.3 m
    cos = np.abs((d32 * n64).sum(1))
    steep = cos >= 0.05
    print(f"[tori] {steep.mean():.4f} of the rays at |cos| >= 0.05", flush=True)
    same = va32 == va64
    assert same[steep].all(), np.flatnonzero(~same & steep)[:8]
    err = np.abs(t32 - t64)
    bound = 1e-5 * t64 + 2e-6
    assert (err[steep] <= bound[steep]).all(), (err[steep] / np.maximum(t64[steep], 1e-9)).max()
    graze = ~steep & same
    assert (err[graze] <= 4.0 * np.sqrt(1.19e-7 * t64[graze] * 0.3) + bound[graze]).all()
    assert same[~steep].mean() >= 0.9


def test_trace_with_tori_against_oracle(gpu, tmp_path):
    """A point source among the tori (two of glass, one hollow and cut in phi), traced with the torus kernel: the
    photons' fates and vertices against the oracle's for the same photons (the oracle passes the torus face a photon
    sits on within 1e-7 m: a grazing torus root is conditioned to sqrt(eps))."""
    from test_oracle import _tori

    det = _tori(tmp_path)
    src = L.sources.Point((0.0, 0.0, 0.0), lam=420.0, n_photons=1.0)
    n = 1 << 16
    g = L.simulate(det, src, samples=n, seed=3)
    assert engine(det).last_launch()["features"] & 16
    _, h = O.OracleRun(det).trace(src, n, seed=3, bvh=True)
    gh = np.asarray(g.health)
    np.testing.assert_allclose(gh[:4], h[:4], rtol=0, atol=2e-3)  # absorbed in bulk, at surfaces, lost, escaped
    assert gh[9] <= h[9] + 1e-4 and gh[5] <= h[5] + 1e-4  # stuck-photon pushes, abandoned photons: rare in both
    np.testing.assert_allclose(gh[12], h[12], rtol=1e-2)  # vertices


def test_gate_and_events_arguments(gpu, tank1):
    """``simulate(gate=, events=)`` (the reference's api.py:876-878): a gate is the time histogram over it, the same
    light as the detector timed on those nodes; ``events`` gives the outputs a leading event axis of that length."""
    n = 1 << 17
    a = L.simulate(tank1, SRC1, samples=n, seed=3, gate=(0.0, 12.0), gate_step=0.5)
    b = L.simulate(tank1.timed(0.0, 0.5, 25), SRC1, samples=n, seed=3)
    np.testing.assert_allclose(a.time.cpu().numpy(), b.time.cpu().numpy(), rtol=1e-9, atol=1e-12)
    assert a.time.shape[-1] == 25 and a.provenance["time_grid"] == (0.0, 0.5, 25)
    c = L.simulate(tank1, SRC1, samples=n, seed=3, events=3)
    assert tuple(c.arrivals.shape) == (3, 8)
    np.testing.assert_allclose(c.arrivals[0].cpu().numpy(), b.arrivals.cpu().numpy(), rtol=1e-9, atol=1e-12)
    assert float(c.arrivals[1:].abs().sum()) == 0.0
    assert c.events == 3 and b.events == 0  # the reference's Light.event(e): event e's light
    np.testing.assert_array_equal(c.event(0).arrivals.cpu().numpy(), c.arrivals[0].cpu().numpy())
    batch = L.sources.Batch((SRC1, SRC1), (n, n), events=(2, 0))
    d = L.simulate(tank1, batch, seed=3, events=4)
    assert tuple(d.arrivals.shape) == (4, 8) and float(d.arrivals[1].abs().sum()) == 0.0
    with pytest.raises(ValueError, match="events"):
        L.simulate(tank1, batch, seed=3, events=2)


def test_chunk_records_and_groups_arguments(gpu, tank1):
    """``simulate(chunk=, records=)`` and ``Light.group`` (the reference's api.py:876-878, 557-562): a run split
    into launches of ``chunk`` photons traces the same photons (ids fixed) into the same tallies, to the order of
    the FP64 sums; ``records`` changes nothing; the group view is the light on the group's sensors."""
    n = (1 << 17) + 123
    a = L.simulate(tank1, SRC1, samples=n, seed=5)
    b = L.simulate(tank1, SRC1, samples=n, seed=5, chunk=(1 << 15) + 7, records=16)
    assert b.provenance["launches"] == 5 and a.provenance["launches"] == 1
    for k in ("arrivals", "arrival2", "mean_arrival_time", "charge"):
        np.testing.assert_allclose(getattr(b, k).cpu().numpy(), getattr(a, k).cpu().numpy(), rtol=1e-12, atol=0)
    ha, hb = np.asarray(a.health), np.asarray(b.health)
    # the fates' weights are summed per wavefront in float32 before their FP64 cells: another split, other roundings
    np.testing.assert_allclose(hb, ha, rtol=1e-6, atol=1e-9)
    assert hb[13] == n and (hb[[10, 12]] == ha[[10, 12]]).all() and b.kernel_ms > 0  # the counts: exact
    batch = L.sources.Batch((SRC1, SRC1, SRC1), (5000, 0, 70000), events=(1, 0, 1))
    c = L.simulate(tank1, batch, seed=2, events=3)
    d = L.simulate(tank1, batch, seed=2, events=3, chunk=4096)
    np.testing.assert_allclose(d.arrivals.cpu().numpy(), c.arrivals.cpu().numpy(), rtol=1e-12, atol=0)
    np.testing.assert_allclose(d.arrivals_error.cpu().numpy(), c.arrivals_error.cpu().numpy(), rtol=1e-9, atol=0)
    g = a.group("pmt")
    np.testing.assert_array_equal(g.arrivals.cpu().numpy(), a.arrivals.cpu().numpy()[tank1.groups[0].sensors])
    np.testing.assert_allclose(L.readout.pe(g, tank1.params()).cpu().numpy(),
                               L.readout.pe(a, tank1.params(), tank1).cpu().numpy(), rtol=1e-15)
    with pytest.raises(ValueError, match="backend"):
        L.simulate(tank1, SRC1, samples=16, seed=5, backend="xla")
