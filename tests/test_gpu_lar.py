"""The wavelength shifter and the LAr box on the device, through the C ABI: against the oracle (the same photons),
the WLS closed form, the batched voxel sources and the adjoint of a two-channel medium."""

import numpy as np
import pytest
import torch
from scipy import stats

import lumen_amd as L
from lumen_amd import units
from lumen_amd.autograd import simulate_diff, table_tensor
from lumen_amd.batch import BatchRun
from lumen_amd.detector import lar_box, voxel_points
from oracle import oracle as O
from toys import sphere_toy

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lar():
    return lar_box()


def test_wls_enclosing_sensor_counts_the_yield(gpu):
    a, n_idx, tau = 1.0, 1.33, 1.5
    det = sphere_toy(a=a, R=0.0, L_abs=None, L_wls=0.01, wls_yield=0.7, wls_tau=tau, wls_profile="exponential",
                     n=n_idx)
    n = 1 << 18
    light = L.simulate(det, L.sources.Point((0, 0, 0), lam=300.0, n_photons=1.0), samples=n, seed=3)
    s = float(light.arrivals.sum())
    assert abs(s - np.float32(0.7)) < 2e-6  # float32 weights summed in FP64
    v = units.C_VACUUM / n_idx
    mean_t = float(light.mean_arrival_time[0])
    assert abs(mean_t - (a / v + 0.01 / v + tau)) < 0.005 * tau + 0.002 * a / v


def test_lar_box_same_photons_as_oracle(gpu, lar):
    """One 128 nm point: the engine (float32) and the oracle (FP64) trace the same photons; the PMTs' arrivals
    agree within their Monte Carlo errors (chi2 of the pulls), the total within 1 %."""
    src = L.sources.Point((0.3, -0.5, 0.2), lam=128.0, n_photons=1.0)
    n = 1 << 16
    g = L.simulate(lar, src, samples=n, seed=7)
    tal, h = O.OracleRun(lar).trace(src, n, seed=7, bvh=True)
    o = O.light_of(tal, h, lar.n_sensors, 1, n)
    ga, ge = g.arrivals.cpu().numpy(), g.arrivals_error.cpu().numpy()
    assert ga.sum() > 0
    assert abs(ga.sum() / o["arrivals"].sum() - 1) < 0.01
    err = np.hypot(ge, o["arrivals_error"])
    m = err > 0
    pulls = (ga[m] - o["arrivals"][m]) / err[m]
    assert stats.chi2.sf(np.sum(pulls**2), m.sum()) > 1e-3
    assert abs(g.health.balance()) < 1e-3 * g.health.start_weight
    # the moment: a photon deposits at more PMTs than its 4 moment slots hold only after many bounces, and then
    # the oldest slot's square is taken early (its later cross terms are dropped); against the oracle's exact moment
    # of the same photons the summed arrival2 stays within 1 %
    ga2, oa2 = g.arrival2.cpu().numpy(), o["arrival2"]
    print(f"[lar] moment spills {g.health.moment_spills:.0f} of {n} photons; arrival2 {ga2.sum():.6g} against the "
          f"oracle's {oa2.sum():.6g}", flush=True)
    assert abs(ga2.sum() / oa2.sum() - 1.0) < 0.01


def test_lar_box_converts_and_reflects(gpu, lar):
    """The light reaching the PMTs was shifted: every deposit's wavelength is in the emission band, so the mean
    time per arrival exceeds the direct flight, and the health balances."""
    src = L.sources.Point((0.0, 0.0, 0.0), lam=128.0, n_photons=1.0)
    g = L.simulate(lar, src, samples=1 << 18, seed=8)
    h = g.health
    assert h.absorbed_bulk > 0.1  # the shifter's unconverted half and LAr absorption
    assert h.absorbed_surface > 0.01  # the PTFE's 5 %
    assert abs(h.balance()) < 1e-3 * h.start_weight
    t = g.mean_arrival_time.cpu().numpy()
    assert np.all(t[g.arrivals.cpu().numpy() > 0] > 1.9 / (units.C_VACUUM / 1.38))


def test_lar_voxel_batch(gpu, lar):
    """64 of the voxel sources in one launch against one lh_trace each."""
    pts = voxel_points(n=16)[::64]
    batch = L.sources.Batch(tuple(pts), tuple([1 << 12] * len(pts)))
    run = BatchRun(lar, batch, seed=9, graph=True, tune=False)
    one = run.forward().cpu().numpy()
    seq = run.sequential().cpu().numpy()
    np.testing.assert_allclose(one, seq, rtol=1e-9, atol=1e-12)
    assert np.all(one[:, :lar.n_sensors].sum(1) > 0)


def _crn_fd(det, src, key, base, samples, seed, h):
    p0 = det.params()

    def tot(scale):
        v = base * scale
        p = p0.set(key, L.params.Table(p0[key].energy, v))
        return float(L.simulate(det, src, p, samples=samples, seed=seed, reference=p0).arrivals.sum())

    return (tot(1 + h) - tot(1 - h)) / (2 * h)


def test_two_channel_medium_adjoint(gpu):
    """A medium with Rayleigh and a shifter: d(total arrivals)/d(scale) of each length by the adjoint against
    common-random-number central differences (the channel choice and both rates' path terms)."""
    det = sphere_toy(a=1.0, R=0.3, L_abs=5.0, L_ray=0.5, L_wls=0.8, wls_yield=0.7)
    src = L.sources.Point((0.1, 0.0, 0.0), lam=300.0, n_photons=1.0)
    p0 = det.params()
    keys = ["materials/Water/RAYLEIGH", "materials/Water/WLSABSLENGTH"]
    ts = {k: table_tensor(p0, k) for k in keys}
    n = 1 << 20
    light = simulate_diff(det, src, ts, samples=n, seed=10)
    light.arrivals.sum().backward()
    for k in keys:
        base = np.asarray(p0[k].value, np.float64)
        g = float((ts[k].grad.cpu().numpy() * base).sum())
        fd = _crn_fd(det, src, k, base, n, 10, 1e-3)
        assert abs(g - fd) <= 0.01 * abs(fd) + 1e-4, (k, g, fd)


def test_lar_adjoint_rayleigh_and_foil(gpu, lar):
    """The LAr box's adjoint parameters (Rayleigh length, foil reflectivity) against CRN central differences."""
    src = L.sources.Point((0.5, 0.0, -0.4), lam=128.0, n_photons=1.0)
    p0 = lar.params()
    n = 1 << 18
    tr = table_tensor(p0, "materials/LAr/RAYLEIGH")
    tf = table_tensor(p0, "surfaces/PTFE/REFLECTIVITY")
    light = simulate_diff(lar, src, {"materials/LAr/RAYLEIGH": tr, "surfaces/PTFE/REFLECTIVITY": tf}, samples=n,
                          seed=11)
    light.arrivals.sum().backward()
    g_ray = float((tr.grad.cpu().numpy() * np.asarray(p0["materials/LAr/RAYLEIGH"].value)).sum())
    fd_ray = _crn_fd(lar, src, "materials/LAr/RAYLEIGH", np.asarray(p0["materials/LAr/RAYLEIGH"].value), n, 11, 1e-3)
    assert abs(g_ray - fd_ray) <= 0.02 * abs(fd_ray) + 1e-5, (g_ray, fd_ray)
    r0 = float(p0["surfaces/PTFE/REFLECTIVITY"])
    h = 1e-3

    def tot(r):
        return float(L.simulate(lar, src, p0.set("surfaces/PTFE/REFLECTIVITY", r), samples=n, seed=11,
                                reference=p0).arrivals.sum())

    fd_f = (tot(r0 * (1 + h)) - tot(r0 * (1 - h))) / (2 * h * r0)
    g_f = float(tf.grad.cpu().numpy()[0])
    assert abs(g_f - fd_f) <= 0.02 * abs(fd_f) + 1e-5, (g_f, fd_f)
    assert torch.isfinite(tr.grad).all()
