"""The tank's Poisson-likelihood gradient (forward plus adjoint over ~1,034 parameters) against common-random-number
central finite differences at h = 1e-3 on a seeded subset of 32 parameters (the reference's finite-difference method,
compare/gradients.py:1-40): every component within 2 % (plus a floor of 1e-4 of the largest component for those
near zero)."""

import numpy as np
import pytest

import lumen_amd as L
from lumen_amd.detector import muon_track, water_tank
from lumen_amd.gradcheck import TankLikelihood, crn_central

pytestmark = pytest.mark.gpu


def test_poisson_nll_gradient_against_crn_finite_differences(gpu):
    det = water_tank()
    x0, d = muon_track()
    src = L.sources.CherenkovSegment("Water", start=tuple(x0), direction=tuple(d), length=1.0)
    prob = TankLikelihood.build(det, src, target_samples=1 << 22)
    n, seed = 1 << 22, 21
    grads, _ = prob.adjoint(samples=n, seed=seed)
    picks = prob.picks(32, seed=4)
    fd = crn_central(prob.loss, det, src, prob.params, picks, samples=n, seed=seed, h=1e-3)
    g = np.array([grads[k][i] for k, i in picks])
    floor = 1e-4 * max(np.abs(np.concatenate([v.ravel() for v in grads.values()])).max(), 1e-30)
    bad = np.abs(g - fd) > 0.02 * np.abs(fd) + floor
    for (k, i), a, b in zip(picks, g, fd):
        print(f"[c3] {k}[{i}]: adjoint {a:.6g}, finite differences {b:.6g}", flush=True)
    assert not bad.any(), [(picks[j], g[j], fd[j]) for j in np.flatnonzero(bad)]
    assert sum(1 for k, _ in picks if not k.startswith("sensors/")) >= 1  # a tracking parameter is among them
