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
    """Return expected photoelectrons (S,) (or (E, S) for a batched Light) at ``params``."""
    import torch

    charge = light.charge
    dev, dt = charge.device, charge.dtype
    kl, kt = getattr(light, "charge_axes", None) or (charge.shape[-1], 1)
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


# The time histogram's cubic Lagrange split of a unit mass at u in [0, 1) past node j onto nodes j - 1 .. j + 2
# (the kernel's deposit weights), as polynomial coefficients in u, lowest power first, and their antiderivatives.
_SPLIT = {-1: np.array([0.0, -2.0, 3.0, -1.0]) / 6.0, 0: np.array([2.0, -1.0, -2.0, 1.0]) / 2.0,
          1: np.array([0.0, 2.0, 1.0, -1.0]) / 2.0, 2: np.array([0.0, -1.0, 0.0, 1.0]) / 6.0}  # fmt: skip
_SPLIT_INT = {o: np.concatenate([[0.0], c / np.arange(1.0, 5.0)]) for o, c in _SPLIT.items()}


def delay_taps(edges, dt: float):
    """Return (off0, M) for a delay table with bin ``edges`` (ns, non-decreasing, >= 0) on nodes of spacing ``dt``:
    M (taps, bins) float64, M[k, b] the mass a delay uniform on bin b moves a unit node mass onto the node off0 + k
    further on (the reference's ``sensors/components.py:231-258`` ``delay_taps``).  Each entry is the exact integral
    over the bin of the histogram's own cubic split weight of that node, over the bin's width; a zero-width bin is a
    point mass (the split weights at its point), so the delay is as exact as the node grid."""  # noqa: D205, D209
    e = np.asarray(edges, np.float64) / float(dt)
    if e.ndim != 1 or len(e) < 2 or np.any(np.diff(e) < 0) or not np.all(np.isfinite(e)):
        raise ValueError("delay_taps: edges must be a finite, non-decreasing list of at least two values")
    lo, hi = int(np.floor(e.min())) - 1, int(np.floor(e.max())) + 2
    M = np.zeros((hi - lo + 1, len(e) - 1))
    for b in range(len(e) - 1):
        xa, xb = e[b], e[b + 1]
        if xb == xa:
            j = int(np.floor(xa))
            for o, c in _SPLIT.items():
                M[j + o - lo, b] += np.polynomial.polynomial.polyval(xa - j, c)
            continue
        for j in range(int(np.floor(xa)), int(np.ceil(xb))):
            g0, g1 = max(xa, j) - j, min(xb, j + 1) - j
            if g1 <= g0:
                continue
            for o, c in _SPLIT_INT.items():
                pv = np.polynomial.polynomial.polyval
                M[j + o - lo, b] += (pv(g1, c) - pv(g0, c)) / (xb - xa)
    return lo, M


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


# ---------------------------------------------------------------------------------------------- simulated events
NEGATIVE_SHARE = 1e-6  # a sensor's negative expected density, as a share of its total, taken as rounding (cleared)


@dataclass(frozen=True)
class SampledEvent:
    """One simulated event (the reference's ``readout/events.py`` ``SampledEvent``): photoelectrons per sensor ``pe``
    (S,), the hit list ``hits`` (dict of arrays ``sensor``, ``t`` ns, ``q`` measured units, ``n_pe`` photoelectrons
    in the window, ``t_first`` its first photoelectron's time) and ``waveforms`` (S, n), sample k at t0 + k dt, or
    None."""  # noqa: D205, D209

    pe: np.ndarray
    hits: dict
    waveforms: Any
    t0: float
    dt: float


def spe_charges(rng, n: int, sigma: float = 0.3, fraction: float = 0.0, tau: float = 1.0) -> np.ndarray:
    """Return n single-photoelectron charges (pe) of the Gaussian-plus-exponential model, mean 1 pe: with probability
    ``fraction`` an exponential of mean ``tau``, else a Gaussian of width ``sigma`` about (1 - fraction tau) /
    (1 - fraction) (the reference's ``sensors/components.py:355-389`` ``GaussExpCharge``)."""  # noqa: D205, D209
    peak = (1.0 - fraction * tau) / (1.0 - fraction)
    expo = rng.random(n) < fraction
    g = rng.normal(peak, sigma, n)
    return np.where(expo, rng.exponential(tau, n), g)


def _windows(sensor, t, q, window: float):
    """Return per integration window (sensor, first time, charge, photoelectrons): a window opens at a sensor's first
    photoelectron not yet in one and holds those within ``window`` ns of it; windows listed by their rank within the
    sensor, then by sensor."""  # noqa: D205, D209
    if not len(t):
        return [np.zeros(0, np.int64), np.zeros(0), np.zeros(0), np.zeros(0, np.int64)]
    order = np.lexsort((t, sensor))
    s, t, q = sensor[order], t[order], q[order]
    rank, first, charge, count, owner = [], [], [], [], []
    k, cur, start = -1, -1, 0.0
    for i in range(len(t)):
        if s[i] != cur:
            cur, k = s[i], -1
            start = -np.inf
        if t[i] > start + window:  # a new window
            k += 1
            start = t[i]
            rank.append(k)
            owner.append(cur)
            first.append(start)
            charge.append(0.0)
            count.append(0)
        charge[-1] += q[i]
        count[-1] += 1
    o = np.lexsort((np.asarray(owner), np.asarray(rank)))
    return [np.asarray(owner, np.int64)[o], np.asarray(first)[o], np.asarray(charge)[o], np.asarray(count, np.int64)[o]]


def electronics(sensor, t, rng, n_sensors: int, gate=None, *, window: float = 0.0, charge=None, threshold=None,
                charge_noise: float = 0.0, charge_scale: float = 1.0, time_walk=None, pulse=None,
                sampling: float | None = None, gain=1.0, baseline: float = 0.0, noise: float = 0.0) -> SampledEvent:
    """Return the hits and waveforms of given photoelectrons (sensor ids, times ``t`` ns) through the front end (the
    reference's ``readout/events.py:62-103`` ``electronics``): SPE charges (``charge`` = dict(sigma, fraction, tau),
    ``spe_charges``); integration windows of ``window`` ns (``_windows``); the threshold's acceptance of the window
    charge Q (``threshold`` = (knots, accept): linear in Q, constant beyond the end knots, clipped to [0, 1]; None:
    all); the measured charge Q + N(0, charge_noise), recorded when positive, times ``charge_scale``; the hit time the
    window's first photoelectron plus the time walk at the measured charge (``time_walk`` = (knots, shift, sigma):
    shift + N(0, sigma), linear in Q); a hit recorded when its window opens inside ``gate`` (None: all, no waveform).
    Waveforms (with ``pulse`` = dict(t_peak, sigma_rise, sigma_fall) and ``sampling`` ns): each photoelectron's pulse
    (``bigauss``) times its charge and ``gain`` (per sensor or one), at the samples over the gate, plus ``baseline``
    and N(0, noise)."""  # noqa: D205, D209
    sensor, t = np.asarray(sensor, np.int64), np.asarray(t, np.float64)
    if len(sensor) and (sensor.min() < 0 or sensor.max() >= n_sensors):
        raise ValueError(f"electronics: sensor ids must lie in [0, {n_sensors})")
    ch = dict(sigma=0.3, fraction=0.0, tau=1.0) | (charge or {})
    g = spe_charges(rng, len(t), float(ch["sigma"]), float(ch["fraction"]), float(ch["tau"]))
    ws, wt, wq, wn = _windows(sensor, t, g, float(window))
    acc = 1.0
    if threshold is not None:
        kn, ac = (np.asarray(x, np.float64) for x in threshold)
        acc = np.clip(np.interp(wq, kn, ac), 0.0, 1.0)
    qm = wq + rng.normal(0.0, float(charge_noise), len(wq))
    keep = (rng.random(len(wq)) < acc) & (qm > 0.0)
    ws, wt, qm, wn = ws[keep], wt[keep], qm[keep], wn[keep]
    first = wt.copy()
    if time_walk is not None:
        kn, sh, sg = (np.asarray(x, np.float64) for x in time_walk)
        wt = wt + np.interp(qm, kn, sh) + rng.normal(0.0, 1.0, len(wt)) * np.interp(qm, kn, sg)
    if gate is not None:  # the gate reads the window's opening photoelectron; the walk may move the time past it
        inside = (first >= gate[0]) & (first <= gate[1])
        ws, wt, qm, wn, first = ws[inside], wt[inside], qm[inside], wn[inside], first[inside]
    hits = dict(sensor=ws, t=wt, q=qm * float(charge_scale), n_pe=wn, t_first=first)
    pe = np.bincount(sensor, minlength=n_sensors).astype(np.float64)
    if pulse is None or gate is None or sampling is None:
        return SampledEvent(pe, hits, None, 0.0, 0.0)
    n = int(np.ceil((gate[1] - gate[0]) / sampling - 1e-9))
    lo = pulse["t_peak"] - PULSE_SPREAD * pulse["sigma_rise"]
    hi = pulse["t_peak"] + PULSE_SPREAD * pulse["sigma_fall"]
    k0 = np.ceil((t + lo - gate[0]) / sampling).astype(np.int64)
    k = k0[:, None] + np.arange(int(np.ceil((hi - lo) / sampling)) + 1)[None, :]
    lag = gate[0] + k * sampling - t[:, None]
    d = lag - pulse["t_peak"]
    amp = np.exp(-0.5 * (d / np.where(d < 0.0, pulse["sigma_rise"], pulse["sigma_fall"])) ** 2)
    gs = np.broadcast_to(np.asarray(gain, np.float64), (n_sensors,))[sensor]
    ok = (k >= 0) & (k < n)
    wf = np.zeros(n_sensors * n)
    np.add.at(wf, (sensor[:, None] * n + k)[ok], (amp * (g * gs)[:, None])[ok])
    wf = wf.reshape(n_sensors, n) + float(baseline)
    return SampledEvent(pe, hits, wf + rng.normal(0.0, float(noise), wf.shape), float(gate[0]), float(sampling))


def sample(light, params: Params, *, seed: int, gate, tts_sigma: float, bin_width: float | None = None,
           dark_rate=0.0, delay=None, det=None, group: int = 0, **frontend) -> SampledEvent:
    """Return one simulated event of the expected light (the reference's ``readout/events.py:125-151`` ``sample``;
    pseudo-data, not differentiable): photoelectron counts Poisson in each bin of ``pe_density`` over ``gate`` (dark
    counts, delay and transit-time spread included; bins of ``bin_width``, default the time histogram's node
    spacing), uniform within the bin, then ``electronics`` with ``frontend``'s settings.  A negative density beyond
    NEGATIVE_SHARE of a sensor's total (more than the cubic split's rounding) is refused."""  # noqa: D205, D209
    tg = light.provenance.get("time_grid") if light.provenance else None
    if tg is None:
        raise ValueError("sample: the Light does not carry its time grid")
    bw = float(tg[1] if bin_width is None else bin_width)
    rho = pe_density(light, params, gate, bw, tts_sigma, det, group, dark_rate, delay)
    rho = rho.detach().cpu().numpy().astype(np.float64)
    if rho.ndim != 2:
        raise ValueError("sample: one event's light (S sensors): index a batched Light's event first")
    neg, tot = -np.minimum(rho, 0.0).sum(-1), np.abs(rho).sum(-1)
    bad = np.flatnonzero(neg > NEGATIVE_SHARE * tot)
    if bad.size:
        raise ValueError(f"sample: the expected density of sensor(s) {bad[:8].tolist()} is negative beyond "
                         f"{NEGATIVE_SHARE:g} of its total")
    rng = np.random.default_rng(seed)
    counts = rng.poisson(np.maximum(rho, 0.0))
    s, b = np.nonzero(counts)
    reps = counts[s, b]
    s, b = np.repeat(s, reps), np.repeat(b, reps)
    t = gate[0] + (b + rng.random(len(b))) * bw
    return electronics(s, t, rng, rho.shape[0], gate, **frontend)
