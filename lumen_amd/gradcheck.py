"""Gradient checks: central finite differences with common random numbers, and the tank's Poisson-likelihood problem.

``crn_central`` is the reference's finite-difference method (``compare/gradients.py:1-40``): the light at value
v (1 +- h) traced with the same seed and every random choice drawn at the unperturbed values (``reference=``), so
the difference quotient converges to the derivative of the light at a fixed seed -- the quantity the adjoint
returns.  ``TankLikelihood`` is the water tank's fit problem: the Poisson negative log-likelihood of the expected
photoelectrons against a target simulated at parameters perturbed by N(0, 5 %), over the 16 absorption knots, the 16
Rayleigh knots, the liner's reflectivity and facet width, and the 1,000 PMT QEs (the QEs act after tracking).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .params import Params, Table

TRACKING_KEYS = ("materials/Water/ABSLENGTH", "materials/Water/RAYLEIGH", "surfaces/Liner/REFLECTIVITY",
                 "surfaces/Liner/sigma_alpha")
QE_KEY = "sensors/pmt/qe"


def values_of(params: Params, key) -> np.ndarray:
    """Return a parameter's values as a float64 vector (a table's values, or the number(s))."""
    v = params[key]
    return np.atleast_1d(np.asarray(v.value if isinstance(v, Table) else v, np.float64))


def with_value(params: Params, key, i: int, x: float) -> Params:
    """Return ``params`` with value i of ``key`` set to x."""
    v = values_of(params, key).copy()
    v[i] = x
    return params.set(key, v if v.size > 1 else float(v[0]))


def after_tracking(key: str) -> bool:
    """Whether a key acts after tracking (a sensor value): its light needs no new trace."""
    return key.startswith("sensors/")


def crn_central(loss, det, source, params: Params, picks, *, samples: int, seed: int, h: float = 1e-3,
                controls=None, device="cuda"):
    """Return the central differences d loss / d value for each (key, index) of ``picks``: (f(v (1 + h)) -
    f(v (1 - h))) / (2 h v), f = loss(light, values), each light traced at the same seed with its random choices
    drawn at ``params`` (common random numbers).  A key acting after tracking reuses the unperturbed light."""  # noqa: D205, D209
    from .api import simulate

    base = None
    out = np.zeros(len(picks))
    for j, (key, i) in enumerate(picks):
        v = float(values_of(params, key)[i])
        f = []
        for s in (1.0 + h, 1.0 - h):
            p = with_value(params, key, i, v * s)
            if after_tracking(key):
                if base is None:
                    base = simulate(det, source, params, samples=samples, seed=seed, controls=controls,
                                    device=device)
                light = base
            else:
                light = simulate(det, source, p, samples=samples, seed=seed, reference=params, controls=controls,
                                 device=device)
            f.append(float(loss(light, p)))
        out[j] = (f[0] - f[1]) / (2.0 * h * v)
    return out


@dataclass
class TankLikelihood:
    """The water tank's Poisson likelihood problem (module docstring): ``target`` (1000,) expected photoelectrons
    simulated at the perturbed parameters, ``params`` the values the gradient is taken at."""

    det: object
    source: object
    params: Params
    target: object  # torch float64 (S,)

    @staticmethod
    def build(det, source, *, target_samples: int = 1 << 24, seed: int = 3, device="cuda") -> "TankLikelihood":
        """Return the problem: the target simulated at ``params`` x (1 + N(0, 0.05)) (seed ``seed``) with
        ``target_samples`` photons."""  # noqa: D205, D209
        from .api import simulate
        from .readout import pe

        p0 = det.params()
        rng = np.random.default_rng(seed)
        pt = p0
        for k in TRACKING_KEYS + (QE_KEY,):
            arr = values_of(p0, k)
            new = arr * (1.0 + 0.05 * rng.normal(size=arr.shape))
            pt = pt.set(k, new if arr.size > 1 else float(new[0]))
        light = simulate(det, source, pt, samples=target_samples, seed=77, device=device)
        return TankLikelihood(det, source, p0, pe(light, pt, det).detach())

    def nll(self, pe_):
        """The Poisson negative log-likelihood sum_s pe_s - target_s log pe_s (torch)."""
        import torch

        return (pe_ - self.target * torch.log(torch.clamp(pe_, min=1e-30))).sum()

    def loss(self, light, params: Params):
        """The loss of a light at values ``params`` (its QEs applied after tracking)."""
        from .readout import pe

        return self.nll(pe(light, params, self.det))

    def keys(self):
        return TRACKING_KEYS + (QE_KEY,)

    def picks(self, n: int = 32, seed: int = 4, tracking: int = 24):
        """Return n (key, index) parameters drawn without replacement (seed ``seed``): ``tracking`` of them from the
        34 that the photon loop reads and the rest from the QEs (a uniform draw over all 1,034 would pick almost
        only QEs, whose derivative is the readout's alone)."""  # noqa: D205, D209
        track = [(k, i) for k in TRACKING_KEYS for i in range(len(values_of(self.params, k)))]
        qe = [(QE_KEY, i) for i in range(len(values_of(self.params, QE_KEY)))]
        rng = np.random.default_rng(seed)
        m = min(tracking, len(track), n)
        a = [track[j] for j in sorted(rng.choice(len(track), m, replace=False))]
        return a + [qe[j] for j in sorted(rng.choice(len(qe), n - m, replace=False))]

    def adjoint(self, *, samples: int, seed: int, controls=None, to_host: bool = True):
        """Return ({key: gradient (values,)}, the light) of the loss by the adjoint (tracking keys) and torch
        autograd (the QEs); ``to_host`` False: the gradients as device tensors, nothing waited for."""  # noqa: D205, D209
        import torch

        from .autograd import simulate_diff, table_tensor
        from .readout import pe

        ts = {k: table_tensor(self.params, k) for k in TRACKING_KEYS}
        q = torch.as_tensor(values_of(self.params, QE_KEY), device=self.target.device).requires_grad_(True)
        light = simulate_diff(self.det, self.source, ts, samples=samples, seed=seed, controls=controls)
        # readout.pe with the QEs as a tensor: efficiency x QE x the sensor's one charge cell
        one = pe(light, self.params.set(QE_KEY, np.ones_like(values_of(self.params, QE_KEY))), self.det)
        loss = self.nll(one * q)
        loss.backward()
        grads = {k: t.grad.detach() for k, t in ts.items()}
        grads[QE_KEY] = q.grad.detach()
        if to_host:
            grads = {k: v.cpu().numpy().astype(np.float64) for k, v in grads.items()}
        return grads, light
