"""Derivatives through the photon loop: ``torch.autograd.Function`` whose backward is the HIP adjoint.

    params_t = {"materials/Water/ABSLENGTH": torch.tensor(..., requires_grad=True), ...}
    light = simulate_diff(det, source, params_t, samples=2**24, seed=0)
    loss = f(light.arrivals, light.charge)            # any torch function of the tallies
    loss.backward()                                   # lh_adjoint: one replayed walk pair per photon

The keys' values are scattered into the packed values (``lumen_amd.pack``) by differentiable torch indexing; the
photon loop's forward is ``lh_trace`` and its backward ``lh_adjoint`` (the photons replayed from their Philox
counters, no stored paths).  Every random choice is drawn at the reference values (default: the values given,
detached), as the reference's derivative is (``api.py:748-787``, ``transport/estimator.py``): the gradient is the
derivative of the light at a fixed seed, which common-random-number finite differences converge to.  Sensor values
(QE, efficiency) act after tracking: their derivatives are ordinary torch autograd on the readout.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib, pack as packing
from .api import Controls, Light, _lazy, engine, light_of, source_rows
from .params import Params, Table


@dataclass
class _Run:
    eng: object
    rows: object
    n_ev: int
    seed: int
    controls: Controls
    ref_values: object  # torch tensor or None
    slots: object  # torch int32 tensor
    n_grad: int
    value_index: object  # torch long: the packed values that carry a slot
    slot_of_value: object  # torch long: their slots
    first_index: object = None  # torch long: the first packed value of each slot (a key read by several tables
    first_slot: object = None  # is packed once per table: its gradient goes to one of its copies)
    health: object = None
    adj_health: object = None
    kernel_ms: float = float("nan")
    adjoint_ms: float = float("nan")


# the launches' kernel times: read on first use (the launches only enqueue)
_Run.kernel_ms = _lazy("kernel_ms", float)
_Run.adjoint_ms = _lazy("adjoint_ms", float)


def _trace(run: _Run, values):
    import torch

    eng = run.eng
    tally = torch.zeros((run.n_ev, eng.stride), dtype=torch.float64, device=eng.device)
    health = torch.zeros(_lib.HEALTH_N, dtype=torch.float64, device=eng.device)
    st = torch.cuda.current_stream(eng.device)
    ctl = run.controls.struct()
    ref = run.ref_values
    _lib.check(_lib.lib().lh_trace(eng.handle, _lib.C.c_void_p(values.data_ptr()),
                                   None if ref is None else _lib.C.c_void_p(ref.data_ptr()), int(values.numel()),
                                   run.rows, len(run.rows), _lib.C.c_uint64(run.seed), _lib.C.byref(ctl),
                                   _lib.C.c_void_p(tally.data_ptr()), run.n_ev, _lib.C.c_void_p(health.data_ptr()),
                                   _lib.C.c_void_p(st.cuda_stream)), "lh_trace")
    run.health = health
    launch = eng.launches() - 1
    run.kernel_ms = lambda: eng.kernel_ms(launch)
    return tally


def adjoint(run: _Run, values, cot):
    """Return the gradient slots (n_grad,) float64 of the loss whose cotangent on the tallies is ``cot``."""
    import torch

    eng = run.eng
    grad = torch.zeros(max(run.n_grad, 1), dtype=torch.float64, device=eng.device)
    health = torch.zeros(_lib.HEALTH_N, dtype=torch.float64, device=eng.device)
    cot = cot.to(torch.float64).contiguous()
    st = torch.cuda.current_stream(eng.device)
    ctl = run.controls.struct()
    ref = run.ref_values
    _lib.check(_lib.lib().lh_adjoint(eng.handle, _lib.C.c_void_p(values.data_ptr()),
                                     None if ref is None else _lib.C.c_void_p(ref.data_ptr()), int(values.numel()),
                                     _lib.C.c_void_p(run.slots.data_ptr()), run.rows, len(run.rows),
                                     _lib.C.c_uint64(run.seed), _lib.C.byref(ctl), _lib.C.c_void_p(cot.data_ptr()),
                                     run.n_ev, _lib.C.c_void_p(grad.data_ptr()), max(run.n_grad, 1),
                                     _lib.C.c_void_p(health.data_ptr()), _lib.C.c_void_p(st.cuda_stream)),
               "lh_adjoint")
    run.adj_health = health
    launch = eng.launches() - 1
    run.adjoint_ms = lambda: eng.kernel_ms(launch)
    return grad[: run.n_grad]


def _function():
    import torch

    class PhotonLoop(torch.autograd.Function):
        """tally = photon_loop(values): forward lh_trace, backward lh_adjoint."""

        @staticmethod
        def forward(ctx, values, run):
            ctx.run = run
            ctx.save_for_backward(values)
            return _trace(run, values.detach().contiguous())

        @staticmethod
        def backward(ctx, g):
            (values,) = ctx.saved_tensors
            run = ctx.run
            gs = adjoint(run, values.detach().contiguous(), g)
            gv = torch.zeros(values.shape, dtype=torch.float64, device=values.device)
            gv[run.first_index] = gs[run.first_slot]  # the copies of a slot share one value of the key
            return gv.to(values.dtype), None

    return PhotonLoop


_PL = None


NO_ADJOINT = ("WLSCOMPONENT", "WLSMEANNUMBERPHOTONS", "WLSTIMECONSTANT")


def simulate_diff(det, source, params_t: dict, *, samples=None, seed=0, params: Params | None = None,
                  reference: Params | None = None, controls: Controls | None = None, first_id=0,
                  edge_rows="structure", device="cuda") -> Light:
    """Return the Light of a run whose tallies are differentiable in ``params_t`` ({key: tensor of the key's table
    values or number}); other values from ``params`` (default ``det.params()``).  The draws are made at
    ``reference`` (default: these values, detached)."""  # noqa: D205, D209
    import torch

    global _PL
    if _PL is None:
        _PL = _function()
    eng = engine(det, device)
    controls = Controls() if controls is None else controls
    if controls.jump:
        eng.ensure_jump()
    p = det.params() if params is None else params
    keys = list(params_t)
    for k in keys:
        if k.endswith("/RINDEX"):
            raise ValueError(f"simulate_diff: {k} is calibrated by value only (the reference raises "
                             "GradientNotAvailable for it)")
        if k.rsplit("/", 1)[-1] in NO_ADJOINT:
            raise ValueError(f"simulate_diff: {k}: the adjoint differentiates the path's rates and surface shares; "
                             "a re-emission's spectrum, yield and delay are calibrated by value (finite differences)")
    pk = packing.pack(det, p, grad_keys=keys, edge_rows=edge_rows)
    values = eng.upload(pk.values)
    slots, value_index, slot_of_value, first_index, first_slot = _slot_tensors(eng, pk.slots)
    # scatter the tensors into the values, differentiably: each key's slots in order are its table's values
    starts = {k: (s0, n) for k, s0, n in pk.grad_keys}
    parts = []
    for k in keys:
        if k not in starts:
            raise KeyError(f"simulate_diff: {k!r} is not read by the photon loop (a sensor value acts after "
                           "tracking: differentiate the readout instead)")
        s0, n = starts[k]
        t = torch.as_tensor(params_t[k], device=eng.device).reshape(-1).to(torch.float32)
        if t.numel() != n:
            raise ValueError(f"simulate_diff: {k} has {n} values, got {t.numel()}")
        parts.append((s0, n, t))
    slot_vals = torch.zeros(pk.n_grad, dtype=torch.float32, device=eng.device)
    for s0, n, t in parts:
        slot_vals = slot_vals.index_put((torch.arange(s0, s0 + n, device=eng.device),), t)
    vals = values.index_put((value_index,), slot_vals[slot_of_value])
    if reference is None:  # the draws at the values themselves: no separate reference operand (the kernel without it)
        ref_values = None
    else:
        ref_values = torch.as_tensor(packing.pack(det, reference, edge_rows=edge_rows).values, device=eng.device)
    rows, n_ev = source_rows(det, source, samples, first_id)
    run = _Run(eng, rows, n_ev, int(seed) & 0xFFFFFFFFFFFFFFFF, controls, ref_values, slots, pk.n_grad,
               value_index, slot_of_value, first_index, first_slot)
    tally = _PL.apply(vals, run)
    health = run.health
    light = light_of(eng, tally, lambda: health.detach().cpu().numpy(),
                     samples if not hasattr(source, "sources") else None, source)
    light.kernel_ms = lambda: run.kernel_ms
    light.group_specs = list(det.groups or [])
    light.provenance = dict(seed=int(seed), samples=samples, detector=det.name, derivative="adjoint (replay)",
                            keys=keys, run=run)
    return light


def _slot_tensors(eng, slots: np.ndarray):
    """Return (slots int32, the packed values' indices that carry a slot, their slots, the first such value of each
    slot, its slot) on the device, the previous ones again when the slots are the same (uploads from pageable memory
    wait for the stream)."""  # noqa: D205, D209
    import torch

    last = getattr(eng, "_slot_upload", None)
    if last is not None and np.array_equal(last[0], slots):
        return last[1]
    idx = np.flatnonzero(slots >= 0)
    _, first = np.unique(slots[idx], return_index=True)
    first = idx[first]
    out = (torch.as_tensor(slots, device=eng.device), torch.as_tensor(idx, device=eng.device, dtype=torch.long),
           torch.as_tensor(slots[idx], device=eng.device, dtype=torch.long),
           torch.as_tensor(first, device=eng.device, dtype=torch.long),
           torch.as_tensor(slots[first], device=eng.device, dtype=torch.long))
    eng._slot_upload = (slots.copy(), out)
    return out


def table_tensor(params: Params, key, requires_grad=True, device="cuda"):
    """Return a key's values as a float64 tensor (a Table's knot values; a number as a 1-vector)."""
    import torch

    v = params[key]
    v = v.value if isinstance(v, Table) else np.atleast_1d(v)
    return torch.tensor(np.asarray(v, np.float64), device=device, requires_grad=requires_grad)
