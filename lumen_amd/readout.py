"""Readout after tracking: expected photoelectrons per sensor (the reference's ``readout/detection.py:59-104``).

``pe[s] = efficiency_s * sum_{k, j} charge[s, k, j] QE_k CE_j``: the charge cells hold each sensor's arrivals split
on the QE table's wavelength knots and the CE table's theta knots (the cathode position angle) by hat weights
(linear between knots), so a monochromatic source reads QE interpolated linearly in wavelength (toy_pmt: 0.245 at
400 nm).  A QE given per sensor (an (S,) array) multiplies the sensor's cells.  Its derivatives in QE, CE and
efficiency are plain torch autograd: they act after tracking.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np

from .params import Params, Table


def pe(light, params: Params, det=None, group: int = 0):
    """Return expected photoelectrons (S,) (or (E, S) for a batched Light) at ``params``: the sensor group's keys
    are those of ``light``'s own group when it is a group's view (``light.group(name)``, the reference's
    ``readout.pe(light.group(name), params)``), else of ``det.groups[group]``."""  # noqa: D205, D209
    import torch

    charge = light.charge
    dev, dt = charge.device, charge.dtype
    kl, kt = getattr(light, "charge_axes", None) or (charge.shape[-1], 1)
    g = getattr(light, "sensor_group", None)  # a group's view (``Light.group(name)``) carries its group
    if g is None:
        g = det.groups[group] if det is not None and det.groups else None
    qe_key = g.qe if g is not None else None
    eff_key = g.efficiency if g is not None else None
    ce_key = getattr(g, "ce", None) if g is not None else None
    if qe_key is None:  # find the sensors' keys by name
        qe_key = next((k for k in params.keys() if k.startswith("sensors/") and k.endswith("/qe")), None)
        eff_key = next((k for k in params.keys() if k.startswith("sensors/") and k.endswith("/efficiency")), None)
    if ce_key is None and kt > 1:
        ce_key = next((k for k in params.keys() if k.startswith("sensors/") and k.endswith("/ce")), None)
    eff = torch.as_tensor(np.asarray(params[eff_key], np.float64), device=dev, dtype=dt) if eff_key else 1.0
    cells = charge.reshape(*charge.shape[:-1], kl, kt)
    if kt > 1 and ce_key is not None:  # CE(theta) on its knots
        ce = torch.as_tensor(np.asarray(params[ce_key].value, np.float64), device=dev, dtype=dt)
        cells = (cells * ce).sum(-1)
    else:
        cells = cells.sum(-1)
    if qe_key is None:
        return eff * cells.sum(-1)
    q = params[qe_key]
    if isinstance(q, Table):
        qe = torch.as_tensor(np.asarray(q.value, np.float64), device=dev, dtype=dt)
        return eff * (cells * qe).sum(-1)
    qe = torch.as_tensor(np.asarray(q, np.float64), device=dev, dtype=dt)
    return eff * qe * cells[..., 0]


def pe_error(light, params: Params, det=None, group: int = 0):
    """Return the Monte Carlo error of ``pe`` per sensor at the values the run was given (the reference's
    ``readout/detection.py`` ``pe_error``): the pe per arrival times the arrivals' error (exact when the sensor's
    acceptance does not vary over the wavelengths it sees: a QE per sensor, or monochromatic light)."""  # noqa: D205, D209
    import torch

    p = pe(light, params, det, group)
    a = light.arrivals
    return torch.where(a > 0, p / torch.where(a > 0, a, torch.ones_like(a)), torch.zeros_like(a)) * light.arrivals_error


# The time histogram splits a point mass at x (in node spacings) onto the four nearest nodes by the weights of
# cubic Lagrange interpolation; node n receives K(x - n) with K the symmetric interpolation kernel on (-2, 2):
#   K(s) = (1 - s^2)(2 - |s|) / 2            |s| <= 1
#   K(s) = -(|s| - 1)(|s| - 2)(|s| - 3) / 6  1 <= |s| <= 2
# Below: K's four unit pieces as polynomials in s, and G(s) = int_{-2}^{s} K, which reaches 1 at s = 2.
def _kernel_pieces():
    P = np.polynomial.Polynomial
    inner = P.fromroots([-1.0, 1.0]) * P([2.0, 1.0]) * -0.5, P.fromroots([-1.0, 1.0]) * P([2.0, -1.0]) * -0.5
    outer = P.fromroots([-1.0, -2.0, -3.0]) / 6.0, P.fromroots([1.0, 2.0, 3.0]) / -6.0
    return [(-2.0, outer[0]), (-1.0, inner[0]), (0.0, inner[1]), (1.0, outer[1])]


_PIECES = _kernel_pieces()
_ANTI = [(s0, p.integ(lbnd=s0)) for s0, p in _PIECES]


def _kernel(s):
    """K(s) in its factored form, so it is exactly 0 or 1 at whole s."""
    u = np.abs(np.asarray(s, np.float64))
    near = 0.5 * (1.0 - u) * (1.0 + u) * (2.0 - u)
    far = (u - 1.0) * (u - 2.0) * (3.0 - u) / 6.0
    return np.where(u <= 1.0, near, np.where(u < 2.0, far, 0.0))


def _kernel_cdf(s):
    s = np.asarray(s, np.float64)
    out = np.zeros_like(s)
    for s0, q in _ANTI:
        out = out + q(np.clip(s, s0, s0 + 1.0))
    return out


def delay_taps(edges, dt: float):
    """Return (off0, M) for a delay table with bin ``edges`` (ns, non-decreasing, >= 0) on nodes of spacing ``dt``:
    M (taps, bins) float64, M[k, b] the mass that a delay uniform on bin b moves from a unit node mass onto the node
    off0 + k further on (the reference's ``sensors/components.py:231-258`` ``delay_taps``).  A delay x (in node
    spacings) puts K(x - n) on node n, K the cubic Lagrange kernel the time histogram splits with; averaged over a
    bin (a, b) that is (G(b - n) - G(a - n)) / (b - a), G the kernel's antiderivative (closed form, piecewise
    quartic), and a zero-width bin is the point mass K(a - n)."""  # noqa: D205, D209
    e = np.asarray(edges, np.float64) / float(dt)
    if e.ndim != 1 or len(e) < 2 or np.any(np.diff(e) < 0) or not np.all(np.isfinite(e)):
        raise ValueError("delay_taps: edges must be a finite, non-decreasing list of at least two values")
    off0 = int(np.floor(e.min())) - 1
    nodes = np.arange(off0, int(np.floor(e.max())) + 3, dtype=np.float64)[:, None]  # K(x - n) > 0 needs |x - n| < 2
    a, b = e[None, :-1], e[None, 1:]
    width = b - a
    avg = (_kernel_cdf(b - nodes) - _kernel_cdf(a - nodes)) / np.where(width > 0, width, 1.0)
    return off0, np.where(width > 0, avg, _kernel(a - nodes))


def delay_apply(nodes, edges, pdf, dt: float):
    """Return the node masses (..., nt) after a delay table (bin ``edges`` ns, probabilities ``pdf`` per bin,
    normalised here; a torch tensor keeps its autograd): each node's mass spread onto the nodes off0 + k further on
    by the taps M pdf; mass moved past the last node (or before the first) is dropped (the reference's
    ``DelayTable.apply``).  A direct sum over the few taps, not an FFT: the taps span the delay's range only."""  # noqa: D205, D209
    import torch

    off0, M = delay_taps(edges, dt)
    w = torch.as_tensor(pdf, dtype=nodes.dtype, device=nodes.device)
    if w.shape[-1] != M.shape[1] or w.ndim != 1:
        raise ValueError(f"delay_apply: pdf needs {M.shape[1]} values, one per bin")
    taps = torch.as_tensor(M, dtype=nodes.dtype, device=nodes.device) @ (w / w.sum())
    nt = nodes.shape[-1]
    out = torch.zeros_like(nodes)
    for k in range(M.shape[0]):
        s = off0 + k  # out[n] += taps[k] nodes[n - s]
        if s >= nt or s <= -nt:
            continue
        if s >= 0:
            out[..., s:] = out[..., s:] + taps[k] * nodes[..., :nt - s]
        else:
            out[..., :nt + s] = out[..., :nt + s] + taps[k] * nodes[..., -s:]
    return out


def delay_mean(edges, pdf):
    """Return the mean delay (ns) of a delay table (bins uniform: their midpoints, weighted by ``pdf``)."""
    import torch

    w = torch.as_tensor(pdf)
    e = torch.as_tensor(np.asarray(edges, np.float64), dtype=w.dtype if w.is_floating_point() else torch.float64)
    w = w.to(e.dtype)
    return (w * 0.5 * (e[1:] + e[:-1])).sum() / w.sum()


def time_masses(light, params: Params, det=None, group: int = 0, delay=None):
    """Return the time histogram as photoelectrons (node masses (S, nt), late (S, 2), early (S, 2); (q, q t) in each
    overflow), scaled per sensor so nodes and overflow sum to ``pe`` (``readout/detection.py:229-254``
    ``block_masses``).  ``delay`` = dict(edges=..., pdf=...): the node masses go through that delay table
    (``delay_apply``) and the overflow cells move by its mean delay."""  # noqa: D205, D209
    import torch

    if light.time is None:
        raise ValueError("time_masses: no time histogram: simulate a timed detector (Detector.timed)")
    H, late, early = light.time, light.late, light.early
    total = H.sum(-1) + late[..., 0] + early[..., 0]
    scale = pe(light, params, det, group) / torch.where(total != 0.0, total, torch.ones_like(total))
    if delay is not None:
        tg = light.provenance.get("time_grid") if light.provenance else None
        if tg is None:
            raise ValueError("time_masses: the Light does not carry its time grid")
        H = delay_apply(H, delay["edges"], delay["pdf"], tg[1])
        shift = delay_mean(delay["edges"], delay["pdf"]).to(device=H.device, dtype=H.dtype)
        late = torch.stack([late[..., 0], late[..., 1] + late[..., 0] * shift], -1)
        early = torch.stack([early[..., 0], early[..., 1] + early[..., 0] * shift], -1)
    return H * scale[..., None], late * scale[..., None], early * scale[..., None]


def pe_density(light, params: Params, gate, bin_width: float, tts_sigma: float, det=None, group: int = 0,
               dark_rate=0.0, delay=None):
    """Return the expected photoelectrons per time bin (S, n_bins) over ``gate`` = (g0, g1) ns, bins of
    ``bin_width``: the node masses (``time_masses``), each spread by a Gaussian transit-time spread of width
    ``tts_sigma`` ns and integrated over the bins (its CDF differences); the overflow masses at their mean times
    (``readout/detection.py`` ``density`` with a Gaussian ``spread``); plus the dark counts, ``dark_rate`` Hz (per
    sensor or one for all) times the bin width.  ``light.provenance`` carries the grid; ``delay``: a delay table
    (``time_masses``) before the spread."""  # noqa: D205, D209
    import math

    import torch

    if not tts_sigma > 0:
        raise ValueError("pe_density: tts_sigma > 0")
    H, late, early = time_masses(light, params, det, group, delay)
    tg = light.provenance.get("time_grid") if light.provenance else None
    if tg is None:
        raise ValueError("pe_density: the Light does not carry its time grid")
    t0, dt, nt = tg
    dev, dtype = H.device, H.dtype
    n_bins = int(math.ceil((gate[1] - gate[0]) / bin_width - 1e-9))
    edges = gate[0] + bin_width * torch.arange(n_bins + 1, device=dev, dtype=dtype)
    nodes = t0 + dt * torch.arange(nt, device=dev, dtype=dtype)

    def spread(times):  # (..., m) -> (..., m, n_bins): the Gaussian's mass in each bin
        z = (edges - times[..., None]) / (tts_sigma * math.sqrt(2.0))
        c = 0.5 * (1.0 + torch.erf(z))
        return c[..., 1:] - c[..., :-1]

    out = H @ spread(nodes)
    for ov in (late, early):
        q, qt = ov[..., 0], ov[..., 1]
        tm = torch.where(q != 0.0, qt / torch.where(q != 0.0, q, torch.ones_like(q)), torch.zeros_like(q))
        out = out + q[..., None] * spread(tm[..., None])[..., 0, :]
    dark = torch.as_tensor(np.asarray(dark_rate, np.float64), device=dev, dtype=dtype)
    if dark.ndim > 0:
        dark = dark[..., None]
    return out + dark * (bin_width * 1e-9)


def bigauss(t, t_peak: float, sigma_rise: float, sigma_fall: float):
    """The single-photoelectron pulse of two half-Gaussians about ``t_peak`` (ns), unit peak (the reference's
    ``sensors/components.py`` ``BiGaussPulse``)."""  # noqa: D205, D209
    import torch

    d = t - t_peak
    s = torch.where(d < 0.0, torch.full_like(d, sigma_rise), torch.full_like(d, sigma_fall))
    return torch.exp(-0.5 * (d / s) ** 2)


@dataclass(frozen=True)
class Waveform:
    """The expected waveform: ``mean`` and ``variance`` (S, n), sample k at t0 + k dt (ns)."""

    mean: Any
    variance: Any
    t0: float
    dt: float


PULSE_SUB = 4  # points per bin averaging the pulse
PULSE_SPREAD = 6.0  # the pulse's support: t_peak - 6 sigma_rise to t_peak + 6 sigma_fall


def expected_waveform(light, params: Params, gate, sampling: float, tts_sigma: float, pulse: dict, *, gain=1.0,
                      charge_moments=(1.0, 1.0), baseline: float = 0.0, noise: float = 0.0, dark_rate=0.0,
                      steps_per_sample: int = 4, det=None, group: int = 0, delay=None) -> Waveform:
    """Return the expected waveform per sensor, sampled every ``sampling`` ns over ``gate`` (the reference's
    ``readout/waveform.py`` ``expected_waveform``):

        E[V_s(t)]   = baseline + gain_s m1 (rho_s * <h>)(t)
        Var[V_s(t)] = gain_s^2 m2 (rho_s * <h^2>)(t) + noise^2      (Campbell's theorem plus white noise)

    rho the photoelectron density per bin (``pe_density`` with the dark counts) on bins of sampling /
    ``steps_per_sample`` ns, <h> the pulse (``pulse`` = dict(t_peak, sigma_rise, sigma_fall), ``bigauss``) averaged
    over each bin, (m1, m2) the single-photoelectron charge's mean and mean square (``charge_moments``), ``delay`` a
    delay table (``time_masses``).  Every step is linear after the time histogram, so torch differentiates it
    exactly.  Fold n is the expectation at g0 + n dt (a photoelectron uniform in bin m is lag (n - m - 1, n - m] bins
    back: the pulse's bin average); sample k is fold k * steps_per_sample, at g0 + k sampling."""  # noqa: D205, D209
    import math

    import torch

    if steps_per_sample < 1 or not sampling > 0:
        raise ValueError("expected_waveform: sampling > 0 and steps_per_sample >= 1")
    dt = sampling / steps_per_sample
    rho = pe_density(light, params, gate, dt, tts_sigma, det, group, dark_rate, delay)  # (S, nb)
    dev, dtype = rho.device, rho.dtype
    lo = pulse["t_peak"] - PULSE_SPREAD * pulse["sigma_rise"]
    hi = pulse["t_peak"] + PULSE_SPREAD * pulse["sigma_fall"]
    j0, j1 = int(math.floor(lo / dt)), int(math.ceil(hi / dt)) + 1
    j = torch.arange(j0, j1 + 1, device=dev, dtype=dtype)
    sub = (torch.arange(PULSE_SUB, device=dev, dtype=dtype) + 0.5) / PULSE_SUB
    t = (j[:, None] - 1.0 + sub[None, :]) * dt  # bin j of the lag: [(j - 1) dt, j dt]
    h = bigauss(t, pulse["t_peak"], pulse["sigma_rise"], pulse["sigma_fall"])
    k1, k2 = h.mean(-1), (h * h).mean(-1)

    def fold(k):  # W_n = sum_j k_j rho_{n - j}, lags j0..j1 (rho zero outside the gate)
        nb = rho.shape[-1]
        out = torch.zeros_like(rho)
        for i, lag in enumerate(range(j0, j1 + 1)):
            if lag >= nb or lag <= -nb:
                continue
            if lag >= 0:
                out[..., lag:] += k[i] * rho[..., :nb - lag]
            else:
                out[..., :nb + lag] += k[i] * rho[..., -lag:]
        return out

    m1, m2 = charge_moments
    g = torch.as_tensor(np.asarray(gain, np.float64), device=dev, dtype=dtype)
    if g.ndim > 0:
        g = g[..., None]
    idx = torch.arange(0, rho.shape[-1], steps_per_sample, device=dev)
    mean = baseline + g * m1 * fold(k1)[..., idx]
    var = g * g * m2 * fold(k2)[..., idx] + noise * noise
    return Waveform(mean, var, float(gate[0]), float(sampling))


def waveform_nll(w: Waveform, data):
    """Return the Gaussian negative log-likelihood of measured waveforms under the expected one (the reference's
    ``readout/waveform.py`` ``waveform_nll``; an approximation at low occupancy)."""  # noqa: D205, D209
    import math

    import torch

    d = torch.as_tensor(data, dtype=w.mean.dtype, device=w.mean.device)
    if d.shape != w.mean.shape:
        raise ValueError(f"waveform_nll: data of shape {tuple(d.shape)}, the expected waveform is {tuple(w.mean.shape)}")
    return 0.5 * torch.sum((d - w.mean) ** 2 / w.variance + torch.log(2.0 * math.pi * w.variance))
