"""THE ORACLE's Python side -- test infrastructure only (tests/, ``__graft_entry__.smoke()``, bench.py's
cpu_baseline).  It loads ``oracle/liblumen_oracle.so`` (built by ``oracle/Makefile``), the FP64 C++/OpenMP
restatement of the reference's photon loop, and runs it on the same packed inputs the engine takes.  The product
(``lumen_amd``) never imports this module.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblumen_oracle.so")
_LIB = None


def build():
    """Compile the oracle (make in oracle/)."""
    r = subprocess.run(["make", "-C", HERE, "-B", "liblumen_oracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stderr[-4000:]}")
    return LIB_PATH


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        from lumen_amd import _lib as E

        L = C.CDLL(LIB_PATH)
        L.oracle_trace.argtypes = [C.POINTER(E.LhGeometry), C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.POINTER(E.LhSource), C.c_int, C.c_uint64, C.POINTER(E.LhControls), C.c_void_p,
                                   C.c_int, C.c_void_p, C.c_int, C.c_int]
        L.oracle_count_ops.argtypes = [C.POINTER(E.LhGeometry), C.c_void_p, C.c_void_p, C.POINTER(E.LhSource),
                                       C.c_int, C.c_uint64, C.POINTER(E.LhControls), C.c_longlong, C.c_void_p]
        L.oracle_query.argtypes = [C.POINTER(E.LhGeometry), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_locate.argtypes = [C.POINTER(E.LhGeometry), C.c_int, C.c_void_p, C.c_void_p]
        L.oracle_philox.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_philox.restype = None
        L.oracle_facet_paths.argtypes = [C.c_double, C.c_void_p, C.c_longlong, C.c_uint64, C.c_void_p]
        L.oracle_facet_paths.restype = None
        L.oracle_lobes.argtypes = [C.c_void_p, C.c_void_p, C.c_longlong, C.c_uint64, C.c_void_p, C.c_void_p]
        L.oracle_lobes.restype = None
        L.oracle_fresnel.argtypes = [C.c_double, C.c_double, C.c_double, C.c_void_p]
        L.oracle_fresnel.restype = None
        L.oracle_wls_emit.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_longlong,
                                      C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_mie.argtypes = [C.c_void_p] + [C.c_double] * 6 + [C.c_longlong, C.c_uint64, C.c_void_p]
        L.oracle_set_jump.argtypes = [C.c_int, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]
        L.oracle_set_jump.restype = None
        L.oracle_mie.restype = None
        L.oracle_mie_pdf.argtypes = [C.c_double, C.c_double, C.c_double, C.c_longlong, C.c_void_p, C.c_void_p]
        L.oracle_mie_pdf.restype = None
        L.oracle_trace_f32.argtypes = L.oracle_trace.argtypes
        L.oracle_glisur_paths.argtypes = [C.c_int, C.c_double, C.c_void_p, C.c_double, C.c_double, C.c_void_p,
                                          C.c_longlong, C.c_uint64, C.c_int, C.c_void_p]
        L.oracle_glisur_paths.restype = None
        L.oracle_glisur_facets.argtypes = [C.c_double, C.c_void_p, C.c_double, C.c_longlong, C.c_uint64, C.c_void_p]
        L.oracle_glisur_facets.restype = None
        L.oracle_fresnel_complex.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_void_p]
        L.oracle_fresnel_complex.restype = None
        L.oracle_emit.argtypes = [C.POINTER(E.LhGeometry), C.c_void_p, C.c_void_p, C.POINTER(E.LhSource), C.c_int,
                                  C.c_uint64, C.c_longlong, C.c_void_p]
        L.oracle_pol_recombine.argtypes = [C.c_void_p] * 6
        L.oracle_pol_recombine.restype = None
        L.oracle_pol_mirror.argtypes = [C.c_void_p] * 3
        L.oracle_pol_mirror.restype = None
        _LIB = L
    return _LIB


def _p(a):
    return C.c_void_p(a.ctypes.data)


class OracleRun:
    """The oracle's view of one detector: its geometry struct and packed layout (kept alive together)."""

    def __init__(self, det):
        from lumen_amd import _lib as E, pack

        self.det = det
        self.geo, self._keep = E.geometry_struct(det.nav, det.sensor_of_vol, det.reach)
        self.S = det.n_sensors

    def trace(self, source, samples=None, *, seed=0, params=None, reference=None, controls=None, threads=0,
              first_id=0, edge_rows="structure", bvh=False, fp32=False):
        """Return (tally (n_events, stride) float64, health (16,)) of the oracle's run.  ``bvh``: find daughters by
        the navigator arrays' BVH (the reference's walk; the CPU baseline), else by brute force (the referee).
        ``fp32``: the photon's state and arithmetic in float32 (the CPU baseline's FP32 figure; not a checker)."""
        from lumen_amd import api, pack

        p = self.det.params() if params is None else params
        pk = pack.pack(self.det, p, edge_rows=edge_rows)
        ref = None if reference is None else pack.pack(self.det, reference, edge_rows=edge_rows).values
        rows, n_ev = api.source_rows(self.det, source, samples, first_id)
        nt = int(pk.layout[11])
        stride = 3 * self.S + self.S * int(pk.layout[5]) * int(pk.layout[14]) + (self.S * (nt + 4) if nt else 0)
        tally = np.zeros((n_ev, stride))
        health = np.zeros(16)
        controls = controls or api.Controls()
        ctl = controls.struct()
        lib().oracle_set_jump(-1, None, 0.0, None, None)
        if controls.jump:  # the same clearance grids the engine reads
            from lumen_amd.detector.clearance import jump_grids

            if getattr(self, "_jump", None) is None:
                self._jump = jump_grids(self.det)
            for g in self._jump:
                lo = np.ascontiguousarray(g.lo, np.float64)
                n = np.ascontiguousarray(g.n, np.int32)
                c = np.ascontiguousarray(g.clear, np.float32)
                lib().oracle_set_jump(int(g.medium), _p(lo), float(g.h), _p(n), _p(c))
        vals = np.ascontiguousarray(pk.values, np.float32)
        lay = np.ascontiguousarray(pk.layout, np.int32)
        refv = None if ref is None else np.ascontiguousarray(ref, np.float32)
        fn = lib().oracle_trace_f32 if fp32 else lib().oracle_trace
        rc = fn(C.byref(self.geo), _p(lay), _p(vals), None if refv is None else _p(refv), rows,
                                len(rows), C.c_uint64(int(seed)), C.byref(ctl), _p(tally), n_ev, _p(health),
                                int(threads), int(bool(bvh)))
        if rc != 0:
            raise RuntimeError(f"oracle_trace: {rc}")
        return tally, health

    def count_ops(self, source, samples, n, *, seed=0, controls=None):
        """Return (FP ops, Philox integer ops, vertices) of the first n photons (the algorithmic work)."""
        from lumen_amd import api, pack

        pk = pack.pack(self.det)
        rows, _ = api.source_rows(self.det, source, samples)
        ctl = (controls or api.Controls()).struct()
        out = np.zeros(3)
        vals = np.ascontiguousarray(pk.values, np.float32)
        lay = np.ascontiguousarray(pk.layout, np.int32)
        lib().oracle_count_ops(C.byref(self.geo), _p(lay), _p(vals), rows, len(rows), C.c_uint64(seed),
                               C.byref(ctl), int(n), _p(out))
        return out

    def emit(self, source, samples, n, *, seed=0, params=None):
        """Return the first n photons as emitted: dict of x (n, 3), d (n, 3), t, lam, w0, er (n, 3), ei (n, 3), vol."""
        from lumen_amd import api, pack

        pk = pack.pack(self.det, self.det.params() if params is None else params)
        rows, _ = api.source_rows(self.det, source, samples)
        out = np.zeros((int(n), 16))
        vals = np.ascontiguousarray(pk.values, np.float32)
        lay = np.ascontiguousarray(pk.layout, np.int32)
        got = lib().oracle_emit(C.byref(self.geo), _p(lay), _p(vals), rows, len(rows), C.c_uint64(int(seed)), int(n),
                                _p(out))
        out = out[:got]
        return dict(x=out[:, 0:3], d=out[:, 3:6], t=out[:, 6], lam=out[:, 7], w0=out[:, 8], er=out[:, 9:12],
                    ei=out[:, 12:15], vol=out[:, 15].astype(np.int64))

    def query(self, x, d, vol, last=None, tmax=1e30):
        """Return (t, vol_after, normal, srf) of rays (FP64 referee)."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(d, np.float64).reshape(-1, 3)
        n = len(x)
        vol = np.ascontiguousarray(np.broadcast_to(np.asarray(vol, np.int32), (n,)))
        last = np.ascontiguousarray(np.full(n, -1, np.int64) if last is None else np.asarray(last, np.int64))
        t, va, nrm, srf = np.zeros(n), np.zeros(n, np.int32), np.zeros((n, 3)), np.zeros(n, np.int64)
        lib().oracle_query(C.byref(self.geo), n, _p(x), _p(d), _p(vol), _p(last), float(tmax), _p(t), _p(va),
                           _p(nrm), _p(srf))
        return t, va, nrm, srf

    def locate(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        out = np.zeros(len(x), np.int32)
        lib().oracle_locate(C.byref(self.geo), len(x), _p(x), _p(out))
        return out


def philox(ctr, key):
    """Return Philox4x32-10 of a counter (4 uint32) and key (2 uint32)."""
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().oracle_philox(_p(c), _p(k), _p(out))
    return out


def facet_paths(sigma, wi, n, seed=0):
    """Return n outgoing directions of the unified facet path off a flat surface (normal +z) at incidence wi."""
    w = np.ascontiguousarray(wi, np.float64)
    out = np.zeros((int(n), 3))
    lib().oracle_facet_paths(float(sigma), _p(w), int(n), C.c_uint64(seed), _p(out))
    return out


def lobes(er, ei, n, seed=0):
    """Return (n Rayleigh directions for the state (er, ei), n Lambertian directions about +z)."""
    a, b = np.ascontiguousarray(er, np.float64), np.ascontiguousarray(ei, np.float64)
    ray, lam = np.zeros((int(n), 3)), np.zeros((int(n), 3))
    lib().oracle_lobes(_p(a), _p(b), int(n), C.c_uint64(seed), _p(ray), _p(lam))
    return ray, lam


def fresnel(c, n1, n2):
    """Return complex (r_s, r_p, t_s, t_p) at incidence cosine c from n1 into n2."""
    out = np.zeros(8)
    lib().oracle_fresnel(float(c), float(n1), float(n2), _p(out))
    return out[0::2] + 1j * out[1::2]


def fresnel_complex(c, n1, nr, nk):
    """Return the oracle's complex amplitudes (r_s, r_p, t_s, t_p) from n1 into the complex index nr + i nk."""
    out = np.zeros(8)
    lib().oracle_fresnel_complex(float(c), float(n1), float(nr), float(nk), _p(out))
    return out[0::2] + 1j * out[1::2]


def wls_emit(packed, medium, lam_in, u_lam, u_time, ref_packed=None):
    """Return (n, 4) re-emissions by ``medium``'s wavelength shifter of photons absorbed at ``lam_in`` nm: the
    wavelength (nm), delay (ns), weight and reference weight, from the uniforms ``u_lam``, ``u_time`` (n each)."""  # noqa: D205, D209
    ul, ut = np.ascontiguousarray(u_lam, np.float64), np.ascontiguousarray(u_time, np.float64)
    out = np.zeros((len(ul), 4))
    ref = None if ref_packed is None else _p(ref_packed.values)
    rc = lib().oracle_wls_emit(_p(packed.layout), _p(packed.values), ref, int(medium), float(lam_in), len(ul), _p(ul),
                               _p(ut), _p(out))
    if rc:
        raise ValueError(f"medium {medium} shifts no wavelengths")
    return out


def mie(wi, gf, gb, r, n, seed=0, live=None):
    """Return (directions (n, 3), weights (n,)) of n Mie draws about ``wi`` at the constants (gf, gb, r), drawn under
    the cover of r, weighted by the density of ``live`` = (gf, gb, r) (default: the drawn constants) over the drawn
    one."""  # noqa: D205, D209
    lf, lb, lr = (gf, gb, r) if live is None else live
    out = np.zeros((int(n), 4))
    w = np.ascontiguousarray(wi, np.float64)
    lib().oracle_mie(_p(w), float(gf), float(gb), float(r), float(lf), float(lb), float(lr), int(n), int(seed), _p(out))
    return out[:, :3], out[:, 3]


def mie_pdf(gf, gb, r, mu):
    """Return the Mie mixture density per steradian at the cosines ``mu``."""
    m = np.ascontiguousarray(mu, np.float64)
    out = np.zeros(len(m))
    lib().oracle_mie_pdf(float(gf), float(gb), float(r), len(m), _p(m), _p(out))
    return out


def light_of(tally, health, S, K, samples):
    """Return dict(arrivals, arrivals_error, mean_arrival_time, arrival2, charge) of an oracle tally (one event)."""
    t = tally[0]
    arr, arr_t, arr2 = t[:S], t[S:2 * S], t[2 * S:3 * S]
    n = float(max(samples, 2))
    err = np.sqrt(np.maximum(n / (n - 1) * (arr2 - arr * arr / n), 0.0))
    mean_t = np.where(arr > 0, arr_t / np.where(arr > 0, arr, 1.0), 0.0)
    out = dict(arrivals=arr, arrivals_error=err, mean_arrival_time=mean_t, arrival2=arr2,
               charge=t[3 * S:3 * S + S * K].reshape(S, K), health=health)
    rest = t[3 * S + S * K:]
    if rest.size:  # a timed detector: the time histogram (S, nt) and its overflow (S, 2 late/early, 2 q/qt)
        nt = rest.size // S - 4
        out["time"] = rest[:S * nt].reshape(S, nt)
        ov = rest[S * nt:].reshape(S, 2, 2)
        out["late"], out["early"] = ov[:, 0], ov[:, 1]
    return out


def pol_recombine(wi, n, a_s, a_p, k, er, ei):
    """Return the oracle's (er, ei) after a specular reflection or refraction: the state (er, ei) along wi at a
    surface of normal n, recombined on the s axis with the complex amplitudes (a_s, a_p) into direction k."""
    v = [np.ascontiguousarray(x, np.float64) for x in (wi, n, [a_s.real, a_s.imag, a_p.real, a_p.imag], k)]
    e = np.ascontiguousarray(np.concatenate([er, ei]), np.float64)
    out = np.zeros(6)
    lib().oracle_pol_recombine(*(_p(x) for x in v), _p(e), _p(out))
    return out[:3], out[3:]


def pol_mirror(h, er, ei):
    """Return the oracle's (er, ei) after a mirror of normal h (the Lambertian and facet paths)."""
    hh = np.ascontiguousarray(h, np.float64)
    e = np.ascontiguousarray(np.concatenate([er, ei]), np.float64)
    out = np.zeros(6)
    lib().oracle_pol_mirror(_p(hh), _p(e), _p(out))
    return out[:3], out[3:]


def glisur_paths(r, wi, n, seed=0, er=None, dielectric=False, n1=1.0, n2=1.0, g4conv=False):
    """Return (last directions (n, 3), far side (n,)) of n GLISUR facet paths off a flat surface (normal +z) at
    incidence wi, width r = 1 - polish: the dielectric's from n1 into n2 or the metal's; ``er`` the states (n, 3)."""
    w = np.ascontiguousarray(wi, np.float64)
    e = np.ascontiguousarray(np.tile([0.0, 1.0, 0.0], (int(n), 1)) if er is None else er, np.float64)
    out = np.zeros((int(n), 4))
    lib().oracle_glisur_paths(int(bool(dielectric)), float(r), _p(w), float(n1), float(n2), _p(e), int(n),
                              C.c_uint64(int(seed)), int(bool(g4conv)), _p(out))
    return out[:, :3], out[:, 3] > 0.5


def glisur_facets(r, V, n, seed=0, e=0.0):
    """Return (facets (n, 3), their log densities) of GLISUR facets drawn for V (local frame) at width r."""
    v = np.ascontiguousarray(V, np.float64)
    out = np.zeros((int(n), 4))
    lib().oracle_glisur_facets(float(r), _p(v), float(e), int(n), C.c_uint64(int(seed)), _p(out))
    return out[:, :3], out[:, 3]
