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
