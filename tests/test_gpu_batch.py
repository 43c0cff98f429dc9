"""The batched launch: 256 sources staged once, traced by one persistent launch over the device work queue, captured
in a HIP graph; against one ``lh_trace`` per source (the same photons: the same rows, ids and seed)."""

import numpy as np
import pytest
import torch

import lumen_amd as L
from lumen_amd.api import Controls, engine
from lumen_amd.batch import BatchRun, c4_batch
from lumen_amd.detector import water_tank

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def det():
    return water_tank()


def _small_batch(det, n=32):
    b = c4_batch(det, n_sources=n)
    counts = tuple(min(c, 1 << 13) for c in b.samples)
    return L.sources.Batch(b.sources, counts)


def test_batched_launch_equals_sequential(gpu, det):
    run = BatchRun(det, _small_batch(det), seed=11, graph=False, tune=False)
    seq = run.sequential().cpu().numpy()
    one = run.forward().cpu().numpy()
    assert np.abs(one).sum() > 0
    np.testing.assert_allclose(one, seq, rtol=1e-9, atol=1e-12)


def test_graph_replays_match_eager(gpu, det):
    """Three replays with eager work and synchronisations between them, each against the eager launch."""
    run = BatchRun(det, _small_batch(det), seed=12, graph=True, tune=False)
    eager = BatchRun(det, _small_batch(det), seed=12, graph=False, tune=False).forward().cpu().numpy()
    for r in range(3):
        out = run.forward().cpu().numpy()
        torch.cuda.synchronize()
        junk = torch.randn(1 << 16, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        assert float(junk) == float(junk)
        np.testing.assert_allclose(out, eager, rtol=1e-9, atol=1e-12, err_msg=f"replay {r}")


def test_graph_forward_adjoint_matches_eager(gpu, det):
    keys = ["materials/Water/ABSLENGTH"]
    qe = torch.as_tensor(det.params()["sensors/pmt/qe"], device="cuda")

    S = len(qe)
    pattern = qe.to(torch.float64).repeat_interleave(engine(det).K)  # built outside the captured graph

    def loss(tally):  # d(total pe)/d(tally): each sensor's qe on its charge cells (sensor-major, K per sensor)
        cot = torch.zeros_like(tally)
        cot[:, 3 * S:] = pattern
        return cot

    g_run = BatchRun(det, _small_batch(det), seed=13, grad_keys=keys, graph=True, tune=False)
    e_run = BatchRun(det, _small_batch(det), seed=13, grad_keys=keys, graph=False, tune=False)
    ge = e_run.forward_adjoint(loss).cpu().numpy().copy()
    assert np.abs(ge).sum() > 0
    for r in range(2):
        gg = g_run.forward_adjoint(loss).cpu().numpy()
        np.testing.assert_allclose(gg, ge, rtol=1e-9, atol=1e-15, err_msg=f"replay {r}")


def test_tally_in_lds_equals_global(gpu, det):
    """The block-local LDS tallies (flushed once per block) and direct global atomics sum the same deposits; the walk
    over LDS-staged BVH nodes and over nodes in memory (another kernel) finds the same crossings."""
    src = L.sources.Point((0.3, -0.2, 0.4), lam=420.0, n_photons=1e5)
    a = L.simulate(det, src, samples=1 << 18, seed=21)
    info = engine(det).last_launch()
    assert info["nodes_in_lds"] == 1 and info["tally_in_lds"] == 1 and info["features"] == 0, info
    for ctl in (Controls(tally_global=True), Controls(nodes_global=True), Controls(tally_global=True, nodes_global=True)):
        b = L.simulate(det, src, samples=1 << 18, seed=21, controls=ctl)
        info = engine(det).last_launch()
        assert info["nodes_in_lds"] == (0 if ctl.nodes_global else 1), info
        assert info["tally_in_lds"] == (0 if ctl.tally_global else 1), info
        assert bool(info["features"] & 8) == bool(ctl.nodes_global), info  # the walk over nodes in memory
        for k in ("arrivals", "arrival2", "charge", "mean_arrival_time"):
            np.testing.assert_allclose(getattr(a, k).cpu().numpy(), getattr(b, k).cpu().numpy(), rtol=1e-9,
                                       atol=1e-12, err_msg=f"{k} {ctl}")
        np.testing.assert_allclose(np.asarray(a.health), np.asarray(b.health), rtol=1e-9)


def test_full_c4_batch_equals_sequential(gpu, det):
    """The whole c4 workload: 256 point sources with 2^12..2^20 photons each (4.8e7 in all), the autotuned,
    graph-captured batched launch against one ``lh_trace`` per source, forward and forward plus adjoint."""
    batch = c4_batch(det)
    assert len(batch.sources) == 256 and min(batch.samples) >= 1 << 12 and max(batch.samples) <= 1 << 20
    print(f"[c4] 256 sources, {sum(batch.samples)} photons", flush=True)
    run = BatchRun(det, batch, seed=11, graph=True, tune=True)
    one = run.forward().cpu().numpy().copy()
    seq = run.sequential().cpu().numpy()
    assert (np.abs(one[:, :run.eng.S]).sum(1) > 0).all()  # every source reaches a sensor
    np.testing.assert_allclose(one, seq, rtol=1e-9, atol=1e-12)
    h = run.health.cpu().numpy()
    assert h[13] == sum(batch.samples)  # every photon traced once


def test_batched_sources_against_the_oracle(gpu, det):
    """Eight of the c4 sources (2^14 photons each) in one batched launch against the oracle's run of the same
    photons: per source, each PMT's arrivals agree within the Monte Carlo error (chi2 of the pulls) and the totals
    within 1 %."""
    from scipy import stats

    from oracle import oracle as O

    b = c4_batch(det, n_sources=8)
    batch = L.sources.Batch(b.sources, (1 << 14,) * 8)
    run = BatchRun(det, batch, seed=5, graph=False, tune=False)
    got = run.forward().cpu().numpy()
    tal, h = O.OracleRun(det).trace(batch, None, seed=5, bvh=True)
    S = run.eng.S
    ok = 0
    for e in range(8):
        g = O.light_of(got[e:e + 1], h, S, 1, 1 << 14)
        o = O.light_of(tal[e:e + 1], h, S, 1, 1 << 14)
        assert abs(g["arrivals"].sum() / o["arrivals"].sum() - 1.0) < 0.01, e
        err = np.hypot(g["arrivals_error"], o["arrivals_error"])
        m = err > 0
        pulls = (g["arrivals"][m] - o["arrivals"][m]) / err[m]
        ok += stats.chi2.sf(np.sum(pulls**2), m.sum()) > 1e-3
    assert ok == 8


def test_sequential_adjoint_equals_the_batched_one(gpu, det):
    """One ``lh_trace`` + ``lh_adjoint`` per source against the batched forward + adjoint (tuned separately): the same
    gradient slots."""
    keys = ["materials/Water/ABSLENGTH", "surfaces/Liner/REFLECTIVITY"]
    run = BatchRun(det, _small_batch(det, 16), seed=17, grad_keys=keys, graph=False, tune=True)
    S = run.eng.S
    target = 1.05 * 0.25 * run.forward()[:, 3 * S:4 * S].clone()

    def loss(tally):
        pe = tally[:, 3 * S:4 * S] * 0.25
        cot = torch.zeros_like(tally)
        cot[:, 3 * S:4 * S] = 0.25 * (1.0 - target / torch.clamp(pe, min=1e-30))
        return cot

    gb = run.forward_adjoint(loss).cpu().numpy().copy()
    assert run.adj_controls is not None and np.abs(gb).sum() > 0
    gs = run.sequential_adjoint(loss).cpu().numpy()
    np.testing.assert_allclose(gs, gb, rtol=1e-8, atol=1e-12 * np.abs(gb).max())
