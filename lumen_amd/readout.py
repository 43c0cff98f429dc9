"""Readout after tracking: expected photoelectrons per sensor (the reference's ``readout/detection.py:59-104``).

``pe[s] = efficiency_s * sum_k charge[s, k] QE_k``: the charge cells hold each sensor's arrivals split on the QE
table's wavelength knots by hat weights (linear in wavelength), so a monochromatic source reads QE interpolated
linearly in wavelength (toy_pmt: 0.245 at 400 nm).  A QE given per sensor (an (S,) array) multiplies the sensor's one
charge cell.  Its derivatives in QE and efficiency are plain torch autograd: they act after tracking.
"""

from __future__ import annotations

import numpy as np

from .params import Params, Table


def pe(light, params: Params, det=None, group: int = 0):
    """Return expected photoelectrons (S,) (or (E, S) for a batched Light) at ``params``."""
    import torch

    charge = light.charge
    dev, dt = charge.device, charge.dtype
    g = det.groups[group] if det is not None and det.groups else None
    qe_key = g.qe if g is not None else None
    eff_key = g.efficiency if g is not None else None
    if qe_key is None:  # find the sensors' keys by name
        qe_key = next((k for k in params.keys() if k.startswith("sensors/") and k.endswith("/qe")), None)
        eff_key = next((k for k in params.keys() if k.startswith("sensors/") and k.endswith("/efficiency")), None)
    eff = torch.as_tensor(np.asarray(params[eff_key], np.float64), device=dev, dtype=dt) if eff_key else 1.0
    if qe_key is None:
        return eff * charge.sum(-1)
    q = params[qe_key]
    if isinstance(q, Table):
        qe = torch.as_tensor(np.asarray(q.value, np.float64), device=dev, dtype=dt)
        return eff * (charge * qe).sum(-1)
    qe = torch.as_tensor(np.asarray(q, np.float64), device=dev, dtype=dt)
    return eff * qe * charge[..., 0]
