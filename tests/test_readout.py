"""The readout after tracking on synthetic histograms (host torch, no GPU): the photoelectron density with dark
counts and the expected waveform (the reference's readout/waveform.py) against quadratures of their definitions."""

import math

import numpy as np
import pytest
import torch

from lumen_amd.api import Health, Light
from lumen_amd.params import Params
from lumen_amd.readout import bigauss, expected_waveform, pe_density, waveform_nll

T0, DT, NT = 0.0, 0.5, 200  # the time histogram's nodes (ns)
N_PE, T_ARR = 50.0, 30.0  # photoelectrons, all at one node time
PULSE = dict(t_peak=10.0, sigma_rise=2.0, sigma_fall=6.0)


def _light():
    H = torch.zeros(1, NT, dtype=torch.float64)
    H[0, int(round((T_ARR - T0) / DT))] = N_PE
    z = torch.zeros(1, 2, dtype=torch.float64)
    q = torch.tensor([N_PE], dtype=torch.float64)
    return Light(q, q * 0, q * 0 + T_ARR, q[:, None].clone(), q * 0, Health(*([0.0] * 15)), 1000, time=H, late=z,
                 early=z.clone(), provenance={"time_grid": (T0, DT, NT)})


def _gauss_conv(f, t, sigma, lo=-60.0, hi=160.0, m=200001):
    """int phi_sigma(tau - T_ARR) f(t - tau) dtau by the trapezoid rule on a fine grid."""
    tau = np.linspace(lo, hi, m)
    w = np.exp(-0.5 * ((tau - T_ARR) / sigma) ** 2) / (sigma * math.sqrt(2 * math.pi))
    return np.array([np.trapezoid(w * f(ti - tau), tau) for ti in t])


def _h(x):
    return bigauss(torch.as_tensor(x, dtype=torch.float64), **PULSE).numpy()


def test_pe_density_mass_and_dark_counts():
    light = _light()
    gate = (0.0, 100.0)
    rho = pe_density(light, Params({}), gate, 0.5, 1.3)
    assert float(rho.sum()) == pytest.approx(N_PE, rel=1e-9)  # all of it inside the gate
    dark = pe_density(light, Params({}), gate, 0.5, 1.3, dark_rate=2e6)  # 2 MHz: 1e-3 pe per 0.5 ns bin
    assert float((dark - rho).sum()) == pytest.approx(2e6 * 100e-9, rel=1e-9)
    centre = float((rho[0] * (torch.arange(200, dtype=torch.float64) + 0.5) * 0.5).sum() / rho.sum())
    assert centre == pytest.approx(T_ARR, abs=1e-6)


@pytest.mark.parametrize("sigma", [2.0, 4.0])
def test_expected_waveform_is_the_pulse_folded_with_the_density(sigma):
    """mean(t) = baseline + g m1 N int phi(tau) h(t - tau) dtau, var(t) = g^2 m2 N int phi h^2 + noise^2, to the
    bins' O(dt^2) (dt = 0.25 ns against spreads of 2 ns and more)."""
    light = _light()
    gain, m1, m2, base, noise = 3.0, 1.1, 1.5, 100.0, 0.7
    w = expected_waveform(light, Params({}), (0.0, 100.0), 1.0, sigma, PULSE, gain=gain, charge_moments=(m1, m2),
                          baseline=base, noise=noise, steps_per_sample=4)
    t = w.t0 + w.dt * np.arange(w.mean.shape[-1])
    mean = base + gain * m1 * N_PE * _gauss_conv(_h, t, sigma)
    var = gain**2 * m2 * N_PE * _gauss_conv(lambda x: _h(x) ** 2, t, sigma) + noise**2
    peak = gain * m1 * N_PE
    assert np.max(np.abs(w.mean[0].numpy() - mean)) < 2e-3 * peak
    assert np.max(np.abs(w.variance[0].numpy() - var)) < 2e-3 * gain**2 * m2 * N_PE
    # the likelihood is smallest at the expectation itself, and grows with a shifted copy
    at = float(waveform_nll(w, w.mean))
    off = float(waveform_nll(w, torch.roll(w.mean, 3, -1)))
    assert off > at


def test_waveform_is_linear_and_differentiable_in_the_histogram():
    """d sum(mean) / d H_j by autograd equals the waveform of a unit mass at node j (the readout is linear)."""
    light = _light()
    H = light.time.clone().requires_grad_(True)
    light.time = H
    light.charge = H.sum(-1, keepdim=True)  # pe follows the histogram (no QE keys: pe = the charge)
    w = expected_waveform(light, Params({}), (0.0, 100.0), 1.0, 2.0, PULSE, gain=2.0, baseline=5.0)
    w.mean.sum().backward()
    j = int(round((T_ARR - T0) / DT)) + 7
    unit = _light()
    unit.time = torch.zeros_like(unit.time)
    unit.time[0, j] = 1.0
    unit.charge = torch.ones(1, 1, dtype=torch.float64)
    wj = expected_waveform(unit, Params({}), (0.0, 100.0), 1.0, 2.0, PULSE, gain=2.0, baseline=0.0)
    assert float(H.grad[0, j]) == pytest.approx(float(wj.mean.sum()), rel=1e-9)
