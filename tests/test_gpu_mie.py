"""Mie scattering (G4OpMieHG) on the device, through the C ABI: the reference's Mie toys (its
tests/physics/test_mie_transport.py) against the oracle (the same photons), the direct light's closed form, Geant4's
ring toy (tests/golden/g4_mie*_rings.npz: 4e6 photons per case, Geant4 10.7.4 and 11.2.2 identical) and the adjoint
in the lobe constants and the Mie length against common-random-number finite differences."""

import json
import math
from pathlib import Path

import numpy as np
import pytest
from scipy import stats

import lumen_amd as L
from lumen_amd import units
from lumen_amd.autograd import simulate_diff, table_tensor
from oracle import oracle as O
from toys import MIE_E, MIE_L, MIE_N, MIE_R, mie_toy

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).parent / "golden"
LAM = units.HC / MIE_E
BEAM = L.sources.Point((0.0, 0.0, 0.0), lam=LAM, n_photons=1.0, direction=(0.0, 0.0, 1.0))


def test_mie_ring_toy_same_photons_as_oracle(gpu):
    """Case A, 36 rings, the pencil beam: the engine (float32) and the oracle (FP64, brute-force navigator) trace
    the same photons; the rings' arrivals and mean times agree within their Monte Carlo errors."""
    det = mie_toy("A", 36)
    n = 1 << 16
    g = L.simulate(det, BEAM, samples=n, seed=5)
    tal, h = O.OracleRun(det).trace(BEAM, n, seed=5)
    o = O.light_of(tal, h, det.n_sensors, 1, n)
    ga, ge = g.arrivals.cpu().numpy(), g.arrivals_error.cpu().numpy()
    # every photon's expected deposits sum to 1 in expectation; the same photons give the same sum
    assert abs(ga.sum() - o["arrivals"].sum()) < 1e-5 and abs(o["arrivals"].sum() - 1.0) < 0.01
    err = np.hypot(ge, o["arrivals_error"])
    m = err > 0
    pulls = (ga[m] - o["arrivals"][m]) / err[m]
    assert stats.chi2.sf(np.sum(pulls**2), m.sum()) > 1e-3, pulls.round(2).tolist()
    assert abs(ga[0] - o["arrivals"][0]) < 1e-3  # the direct beam and the near-forward scatters
    assert abs(g.health.balance()) < 1e-5 * g.health.start_weight


def test_mie_sphere_charge_and_direct_light(gpu):
    """Isotropic light at the centre of the one-ring sphere: the charge is N (no absorption), and the first vertex
    alone (a run cut after one vertex) deposits exp(-R / L) per photon at R n / c, to float32 accuracy."""
    det = mie_toy("A", 1)
    src = L.sources.Point((0.0, 0.0, 0.0), lam=LAM, n_photons=1.0)
    n = 1 << 20
    g = L.simulate(det, src, samples=n, seed=1)
    q, err = float(g.arrivals[0]), float(g.arrivals_error[0])
    assert abs(q - 1.0) < 5 * err and err > 0, (q, err)
    one = L.simulate(det, src, samples=n, seed=1, controls=L.api.Controls(max_vertices=1))
    assert float(one.arrivals[0]) == pytest.approx(math.exp(-MIE_R / MIE_L), rel=2e-5)
    assert float(one.mean_arrival_time[0]) == pytest.approx(MIE_R * MIE_N / units.C_VACUUM, rel=2e-6)


@pytest.mark.parametrize("case", ["A", "B"])
def test_ring_toy_against_geant4(gpu, case):
    """Per-ring charge and mean arrival time of the pencil beam against Geant4's (4e6 photons each): 16 replicas of
    2^18 photons in one batched launch give our errors; chi2 over the 36 rings and the largest pull."""
    man = json.loads((GOLDEN / "g4_mie_rings_manifest.json").read_text())
    with np.load(GOLDEN / f"g4_mie{case}_rings.npz") as z:
        counts, t_sum, t2_sum, launched = z["counts"], z["t_sum"], z["t2_sum"], int(z["launched"])
    assert man[f"g4_11.2.2_mie{case}_rings.npz"]["hits"] == int(counts.sum())
    det = mie_toy(case, 36)
    k, per = 16, 1 << 18
    light = L.simulate(det, L.sources.Batch((BEAM,) * k, (per,) * k), seed=5)
    q = light.arrivals.cpu().numpy()  # (k, 36) per photon
    tw = light.mean_arrival_time.cpu().numpy() * q
    ours, ours_err = q.mean(0), q.std(0, ddof=1) / math.sqrt(k)
    ref, ref_err = counts / launched, np.sqrt(counts) / launched
    z = (ours - ref) / np.sqrt(ours_err**2 + ref_err**2)
    c2 = float(np.sum(z * z))
    assert stats.chi2.sf(c2, 36) > 1e-3 and np.max(np.abs(z)) < 4.5, (case, c2, z.round(2).tolist())
    use = counts > 1000
    t_ref = t_sum[use] / counts[use]
    t_ref_err = np.sqrt(np.maximum(t2_sum[use] / counts[use] - t_ref**2, 0.0) / counts[use])
    t_ours = tw.sum(0)[use] / q.sum(0)[use]
    t_err = (tw[:, use] / np.maximum(q[:, use], 1e-30)).std(0, ddof=1) / math.sqrt(k)
    zt = (t_ours - t_ref) / np.sqrt(t_err**2 + t_ref_err**2)
    assert np.max(np.abs(zt)) < 4.5, (case, zt.round(2).tolist())


def test_mie_adjoint_against_crn_finite_differences(gpu):
    """d(sum_i cos(theta_i) q_i)/d(each of g_f, g_b, r and the MIEHG scale) over the 36 rings of case A: the
    adjoint (the lobe mixture's score at each Mie event, the Mie rate in the path density and the reach) against
    central differences with the draws held at the reference (common random numbers)."""
    det = mie_toy("A", 36)
    p0 = det.params()
    m = "materials/MieLike/"
    keys = [m + "MIEHG_FORWARD", m + "MIEHG_BACKWARD", m + "MIEHG_FORWARD_RATIO", m + "MIEHG"]
    n = 1 << 20
    theta = (np.arange(36) + 0.5) * np.pi / 36
    c = np.cos(theta)
    import torch

    ts = {k: table_tensor(p0, k) for k in keys}
    light = simulate_diff(det, BEAM, ts, samples=n, seed=12)
    (light.arrivals * torch.as_tensor(c, device=light.arrivals.device)).sum().backward()

    def obs(p):
        return float((L.simulate(det, BEAM, p, samples=n, seed=12, reference=p0).arrivals.cpu().numpy() * c).sum())

    h = 1e-3
    for k in keys[:3]:
        v0 = float(p0[k])
        fd = (obs(p0.set(k, v0 + h)) - obs(p0.set(k, v0 - h))) / (2 * h)
        g = float(ts[k].grad.cpu().numpy()[0])
        assert abs(g - fd) <= 0.02 * abs(fd) + 2e-4 and abs(fd) > 5e-3, (k, g, fd)
    base = np.asarray(p0[keys[3]].value, np.float64)
    fd = (obs(p0.set(keys[3], L.params.Table(p0[keys[3]].energy, base * (1 + h))))
          - obs(p0.set(keys[3], L.params.Table(p0[keys[3]].energy, base * (1 - h))))) / (2 * h)
    g = float((ts[keys[3]].grad.cpu().numpy() * base).sum())
    assert abs(g - fd) <= 0.02 * abs(fd) + 2e-4 and abs(fd) > 5e-3, (keys[3], g, fd)
