"""The allocation-free entry point ``lh_trace_static`` against ``lh_trace`` through the C ABI: the same sources,
values and seed give the same tallies and health, on two calls in a row (the second with fewer sources and photons,
so it runs on what the first left in the device's scratch array), and more sources than the scratch holds are
refused before anything is enqueued."""

import ctypes as C

import numpy as np
import pytest

import lumen_amd as L
from lumen_amd import _lib, pack
from lumen_amd.api import engine, source_rows
from lumen_amd.detector import muon_track, water_tank

pytestmark = pytest.mark.gpu


def _run(fn, eng, rows, n_ev, values, seed):
    import torch

    tally = torch.zeros((n_ev, eng.stride), dtype=torch.float64, device=eng.device)
    health = torch.zeros(_lib.HEALTH_N, dtype=torch.float64, device=eng.device)
    st = torch.cuda.current_stream(eng.device)
    ctl = L.api.Controls().struct()
    _lib.check(fn(eng.handle, C.c_void_p(values.data_ptr()), None, int(values.numel()), rows, len(rows),
                  C.c_uint64(seed), C.byref(ctl), C.c_void_p(tally.data_ptr()), n_ev, C.c_void_p(health.data_ptr()),
                  C.c_void_p(st.cuda_stream)), fn.__name__)
    torch.cuda.synchronize()
    return tally.cpu().numpy(), health.cpu().numpy()


def _batch(det, n_src, scale, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1.2, 1.2, (n_src, 3))
    srcs = tuple(L.sources.Point(tuple(p), lam=float(rng.uniform(350.0, 500.0)), n_photons=1e3) for p in pos)
    counts = tuple(int(c) for c in (scale * 2.0 ** rng.uniform(0.0, 4.0, n_src)).astype(np.int64))
    return L.sources.Batch(srcs, counts)


def test_static_entry_point_equals_lh_trace_twice(gpu):
    det = water_tank()
    eng = engine(det, "cuda:0")
    values = eng.upload(pack.pack(det, det.params()).values)
    x0, d = muon_track()
    cases = [("batch of 300 point sources", _batch(det, 300, 2048, 1), None),
             ("one Cherenkov track, 2^18 photons",
              L.sources.CherenkovSegment("Water", start=tuple(x0), direction=tuple(d), length=1.0), 1 << 18),
             ("batch of 17 point sources", _batch(det, 17, 512, 2), None)]
    lib = _lib.lib()
    for k, (what, src, samples) in enumerate(cases):
        print(f"[no-alloc] case {k}: {what}", flush=True)
        rows, n_ev = source_rows(det, src, samples)
        t0, h0 = _run(lib.lh_trace, eng, rows, n_ev, values, 7 + k)
        t1, h1 = _run(lib.lh_trace_static, eng, rows, n_ev, values, 7 + k)
        assert h0[_lib.HEALTH_NAMES.index("photons")] == sum(r.n for r in rows)
        # FP64 tallies added atomically in another order: equal to a few ulps; the health cells are float32 per-lane
        # sums added atomically, equal to float32 rounding
        np.testing.assert_allclose(t1, t0, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(h1, h0, rtol=1e-6, atol=1e-6)
        assert float(t0[:, :eng.S].sum()) > 0.0


def test_static_entry_point_refuses_too_many_sources(gpu):
    det = water_tank()
    eng = engine(det, "cuda:0")
    values = eng.upload(pack.pack(det, det.params()).values)
    src = L.sources.Batch(tuple(L.sources.Point((0.0, 0.0, 0.0), lam=400.0) for _ in range(4097)), (1,) * 4097,
                          events=(0,) * 4097)
    rows, n_ev = source_rows(det, src, None)
    import torch

    tally = torch.zeros((n_ev, eng.stride), dtype=torch.float64, device=eng.device)
    health = torch.zeros(_lib.HEALTH_N, dtype=torch.float64, device=eng.device)
    ctl = L.api.Controls().struct()
    rc = _lib.lib().lh_trace_static(eng.handle, C.c_void_p(values.data_ptr()), None, int(values.numel()), rows,
                                    len(rows), C.c_uint64(1), C.byref(ctl), C.c_void_p(tally.data_ptr()), n_ev,
                                    C.c_void_p(health.data_ptr()), None)
    assert rc == 1 and b"LH_STATIC_MAX_SOURCES" in _lib.lib().lh_last_error()
    torch.cuda.synchronize()
    assert float(health.abs().sum()) == 0.0
