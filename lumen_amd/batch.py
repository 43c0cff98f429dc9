"""Batched sources in one persistent launch, captured in a HIP graph, with an autotuned work queue.

    run = BatchRun(det, Batch(sources, samples), seed=0, grad_keys=[...])   # stages once, captures the graph
    tally = run.forward()             # one graph replay: zero the tallies, the persistent photon loop
    grads = run.forward_adjoint(loss) # forward, the per-event loss's cotangent in torch, the adjoint: one replay

Every source is one tally event; the device work queue hands wavefronts chunks of photons across all sources, so a
launch of many sources with uneven photon counts keeps every lane busy where one launch per source leaves the chip
idle for the small ones.  The graph is replayed with the values, seed and buffers it was captured with (change them
in place: ``run.values``).  The reference runs one ``simulate`` per source (docs/tutorial/first-hour.md, section 2:
"For several sources, call simulate in a loop"); ``sequential`` times exactly that, source by source.
"""

from __future__ import annotations

import numpy as np

from . import _lib, pack as packing
from .api import Controls, engine, source_rows
from .autotune import tuned_controls
from .sources import Batch


class BatchRun:
    """A staged batch: its sources on the device, its packed values, tallies and (optionally) the captured graphs."""

    def __init__(self, det, batch: Batch, params=None, *, seed=0, controls: Controls | None = None, grad_keys=(),
                 graph=True, tune=True, device="cuda"):
        import torch

        self.det, self.batch = det, batch
        self.eng = engine(det, device)
        dev = self.eng.device
        p = det.params() if params is None else params
        self.pk = packing.pack(det, p, grad_keys=list(grad_keys))
        self.values = torch.as_tensor(self.pk.values, device=dev)
        self.slots = torch.as_tensor(self.pk.slots, device=dev)
        self.n_grad = max(self.pk.n_grad, 1)
        self.rows, self.n_ev = source_rows(det, batch, None)
        total = _lib.C.c_int64()
        handle = _lib.C.c_void_p()
        slots = _lib.C.c_void_p(self.slots.data_ptr()) if grad_keys else None
        _lib.check(_lib.lib().lh_batch_create(self.eng.handle, self.rows, len(self.rows), slots,
                                              int(self.values.numel()), self.n_grad, _lib.C.byref(handle),
                                              _lib.C.byref(total)), "lh_batch_create")
        self.handle = handle
        self.total = total.value
        self.has_grad = bool(grad_keys)
        self.seed = int(seed)
        self.tally = torch.zeros((self.n_ev, self.eng.stride), dtype=torch.float64, device=dev)
        self.health = torch.zeros(_lib.HEALTH_N, dtype=torch.float64, device=dev)
        self.cot = torch.zeros_like(self.tally)
        self.grad = torch.zeros(self.n_grad, dtype=torch.float64, device=dev)
        self.adj_health = torch.zeros_like(self.health)
        self.graph = graph
        self.tune = tune and controls is None
        self.controls = controls if controls is not None else (
            tuned_controls(self, "forward") if tune else Controls())
        self.adj_controls = controls  # the adjoint's own: tuned on its first launch (None until then)
        if self.controls.jump:
            self.eng.ensure_jump()
        self._g_fwd = self._g_both = None
        self._loss = None

    def close(self):
        """Release the batch's device buffers (its sources and work queue); the captured graphs go with them."""
        if getattr(self, "handle", None):
            self._g_fwd = self._g_both = None
            _lib.lib().lh_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass

    # ------------------------------------------------------------------------------------------ raw launches
    def _launch(self, ctl=None):
        import torch

        st = torch.cuda.current_stream(self.eng.device)
        c = (ctl or self.controls).struct()
        _lib.check(_lib.lib().lh_batch_trace(self.handle, _lib.C.c_void_p(self.values.data_ptr()), None,
                                             int(self.values.numel()), _lib.C.c_uint64(self.seed), _lib.C.byref(c),
                                             _lib.C.c_void_p(self.tally.data_ptr()), self.n_ev,
                                             _lib.C.c_void_p(self.health.data_ptr()), _lib.C.c_void_p(st.cuda_stream)),
                   "lh_batch_trace")

    def _launch_adjoint(self, ctl=None):
        import torch

        st = torch.cuda.current_stream(self.eng.device)
        c = (ctl or self.adj_controls or Controls()).struct()
        _lib.check(_lib.lib().lh_batch_adjoint(self.handle, _lib.C.c_void_p(self.values.data_ptr()), None,
                                               int(self.values.numel()), _lib.C.c_uint64(self.seed), _lib.C.byref(c),
                                               _lib.C.c_void_p(self.cot.data_ptr()), self.n_ev,
                                               _lib.C.c_void_p(self.grad.data_ptr()), self.n_grad,
                                               _lib.C.c_void_p(self.adj_health.data_ptr()),
                                               _lib.C.c_void_p(st.cuda_stream)), "lh_batch_adjoint")

    def _fwd_body(self, ctl=None):
        self.tally.zero_()  # a fill kernel (graph-safe), not a memset
        self.health.zero_()
        self._launch(ctl)

    def _both_body(self):
        self._fwd_body()
        self.cot.copy_(self._loss(self.tally))
        self.grad.zero_()
        self.adj_health.zero_()
        self._launch_adjoint()

    def _capture(self, body):
        import torch

        s = torch.cuda.Stream(self.eng.device)
        s.wait_stream(torch.cuda.current_stream(self.eng.device))
        with torch.cuda.stream(s):  # warm-up outside the graph (allocations of the loss's temporaries)
            body()
        torch.cuda.current_stream(self.eng.device).wait_stream(s)
        torch.cuda.synchronize(self.eng.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            body()
        return g

    # ------------------------------------------------------------------------------------------ the runs
    def forward(self, ctl=None):
        """Trace every source; returns the tallies (n_events, stride)."""
        if not self.graph or ctl is not None:
            self._fwd_body(ctl)
            return self.tally
        if self._g_fwd is None:
            self._g_fwd = self._capture(self._fwd_body)
        self._g_fwd.replay()
        return self.tally

    def forward_adjoint(self, loss):
        """Forward, ``cot = loss(tally)`` (a torch function returning dLoss/dtally), the adjoint: the gradient
        slots (n_grad,).  Captured as one graph on first use (the loss function is part of it)."""  # noqa: D205, D209
        if self._loss is not loss:
            self._loss, self._g_both = loss, None
        if self.adj_controls is None:  # the adjoint's launch shape: tuned once, with this loss's cotangent
            self._fwd_body()
            self.cot.copy_(loss(self.tally))
            self.adj_controls = tuned_controls(self, "adjoint") if self.tune else Controls()
        if not self.graph:
            self._both_body()
            return self.grad
        if self._g_both is None:
            self._g_both = self._capture(self._both_body)
        self._g_both.replay()
        return self.grad

    def sequential_adjoint(self, loss):
        """The forward plus adjoint one source at a time, as one ``simulate`` and one backward per source do: each
        source's ``lh_trace``, its cotangent ``loss`` of the tallies (the per-event loss reads its own event's row),
        its ``lh_adjoint``; returns the gradient slots summed over the sources."""  # noqa: D205, D209
        import torch

        out = torch.zeros_like(self.tally)
        health = torch.zeros_like(self.health)
        grad = torch.zeros_like(self.grad)
        c = Controls().struct()
        st = torch.cuda.current_stream(self.eng.device)
        vp = _lib.C.c_void_p(self.values.data_ptr())
        for i in range(len(self.rows)):
            one = (_lib.LhSource * 1)()
            one[0] = self.rows[i]
            _lib.check(_lib.lib().lh_trace(self.eng.handle, vp, None, int(self.values.numel()), one, 1,
                                           _lib.C.c_uint64(self.seed), _lib.C.byref(c),
                                           _lib.C.c_void_p(out.data_ptr()), self.n_ev,
                                           _lib.C.c_void_p(health.data_ptr()), _lib.C.c_void_p(st.cuda_stream)),
                       "lh_trace")
            cot = loss(out).to(torch.float64).contiguous()
            _lib.check(_lib.lib().lh_adjoint(self.eng.handle, vp, None, int(self.values.numel()),
                                             _lib.C.c_void_p(self.slots.data_ptr()), one, 1,
                                             _lib.C.c_uint64(self.seed), _lib.C.byref(c),
                                             _lib.C.c_void_p(cot.data_ptr()), self.n_ev,
                                             _lib.C.c_void_p(grad.data_ptr()), self.n_grad,
                                             _lib.C.c_void_p(health.data_ptr()), _lib.C.c_void_p(st.cuda_stream)),
                       "lh_adjoint")
        return grad

    def sequential(self):
        """The same sources one ``lh_trace`` per source, as one ``simulate`` call per source does; the tallies."""
        import torch

        out = torch.zeros_like(self.tally)
        health = torch.zeros_like(self.health)
        c = Controls().struct()
        st = torch.cuda.current_stream(self.eng.device)
        for i in range(len(self.rows)):
            one = (_lib.LhSource * 1)()
            one[0] = self.rows[i]
            _lib.check(_lib.lib().lh_trace(self.eng.handle, _lib.C.c_void_p(self.values.data_ptr()), None,
                                           int(self.values.numel()), one, 1, _lib.C.c_uint64(self.seed),
                                           _lib.C.byref(c), _lib.C.c_void_p(out.data_ptr()), self.n_ev,
                                           _lib.C.c_void_p(health.data_ptr()), _lib.C.c_void_p(st.cuda_stream)),
                       "lh_trace")
        return out


def c4_batch(det, n_sources=256, seed_pos=5, seed_counts=6, seed_lam=7, r_fid=1.8, hz_fid=1.3):
    """The 256-source batch: isotropic monochromatic points, wavelengths uniform in 350-500 nm, positions uniform in
    the fiducial cylinder, photon counts 2^U with U uniform on [12, 20] (seeded)."""  # noqa: D205, D209
    from .sources import Point

    rp = np.random.default_rng(seed_pos)
    r = r_fid * np.sqrt(rp.uniform(size=n_sources))
    ph = rp.uniform(0, 2 * np.pi, n_sources)
    z = rp.uniform(-hz_fid, hz_fid, n_sources)
    counts = np.round(2.0 ** np.random.default_rng(seed_counts).uniform(12, 20, n_sources)).astype(np.int64)
    lam = np.random.default_rng(seed_lam).uniform(350.0, 500.0, n_sources)
    srcs = tuple(Point((float(r[i] * np.cos(ph[i])), float(r[i] * np.sin(ph[i])), float(z[i])), lam=float(lam[i]),
                       n_photons=float(counts[i])) for i in range(n_sources))
    return Batch(srcs, tuple(int(c) for c in counts))
