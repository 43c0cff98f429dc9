"""The ctypes binding of ``liblumen_hip.so`` (the C ABI of ``include/lumen_hip.h``).

The library is built in-tree by ``lumen_amd.build``; importing this module never builds it.  Loading it where it is
missing raises ``EngineNotBuilt``: there is no CPU fallback for the photon loop.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblumen_hip.so")

LH_SRC_POINT, LH_SRC_BEAM, LH_SRC_CHERENKOV = 0, 1, 2
HEALTH_NAMES = ("absorbed_bulk", "absorbed_surface", "lost", "escaped", "distance_cut", "abandoned", "truncated",
                "rouletted", "reweighted", "pushed", "records", "moment_spills", "vertices", "photons",
                "start_weight", "_unused")
HEALTH_N = 16


class EngineNotBuilt(RuntimeError):
    """The HIP engine's library is not built (run ``python -m lumen_amd.build``)."""


class EngineError(RuntimeError):
    """A C ABI call returned an error."""


_f, _i = C.POINTER(C.c_float), C.POINTER(C.c_int32)


class LhGeometry(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("n_vol", "n_nodes", "n_items", "n_groups", "n_leaves", "n_planes", "n_prog",
                                         "n_kinds")] + [
        ("R", _f), ("t", _f), ("kind", _i), ("parent", _i), ("cls", _i), ("sensor", _i), ("vg0", _i), ("vg1", _i),
        ("gR", _f), ("g0", _i), ("g1", _i), ("lo", _f), ("hi", _f), ("skip", _i), ("start", _i), ("count", _i),
        ("items", _i), ("leaf_typ", _i), ("leaf_p0", _i), ("leaf_np", _i), ("leaf_o", _f), ("leaf_q", _f),
        ("leaf_A", _f), ("leaf_b", _f), ("leaf_c", _f), ("plane_n", _f), ("plane_h", _f), ("kind_l0", _i),
        ("kind_nl", _i), ("kind_p0", _i), ("kind_pl", _i), ("prog_op", _i), ("prog_arg", _i), ("reach", C.c_float), ("skin", _i), ("touch", _i),
        ("box_lo", _f), ("box_hi", _f)]


class LhSource(C.Structure):
    _fields_ = [("kind", C.c_int32), ("medium", C.c_int32), ("event", C.c_int32), ("flags", C.c_int32),
                ("pos", C.c_float * 3), ("dir", C.c_float * 3), ("lam", C.c_float), ("lam_hi", C.c_float),
                ("time", C.c_float), ("weight", C.c_float), ("length", C.c_float), ("beta", C.c_float),
                ("charge", C.c_float), ("pad", C.c_float), ("n", C.c_int64), ("first_id", C.c_int64)]


class LhControls(C.Structure):
    _fields_ = [("roulette", C.c_float), ("max_distance", C.c_float), ("max_vertices", C.c_int32),
                ("chunk", C.c_int32), ("blocks_per_cu", C.c_int32), ("tally_global", C.c_int32),
                ("nodes_global", C.c_int32), ("jump", C.c_int32), ("keep_zero_weight", C.c_int32),
                ("reach_bound", C.c_int32)]


_LIB = None

SYMBOLS = ("lh_model_create", "lh_model_destroy", "lh_model_info", "lh_trace", "lh_adjoint", "lh_query",
           "lh_batch_create", "lh_batch_destroy", "lh_batch_trace", "lh_batch_adjoint", "lh_last_kernel_ms",
           "lh_last_launch", "lh_model_jump", "lh_bvh_build", "lh_launches", "lh_kernel_ms_of", "lh_trace_static",
           "lh_last_error")


def lib():
    """Return the loaded library (``EngineNotBuilt`` where it is missing)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise EngineNotBuilt(f"{LIB_PATH} is missing: build it with `python -m lumen_amd.build` (hipcc, gfx950)")
        # torch's HIP runtime first: the library binds to the runtime already in the process, and a process whose
        # first HIP runtime was the library's own found no device on the GPU machine
        import torch  # noqa: F401

        L = C.CDLL(LIB_PATH)
        L.lh_model_create.argtypes = [C.POINTER(LhGeometry), _i, C.c_int32, C.POINTER(C.c_void_p)]
        L.lh_model_destroy.argtypes = [C.c_void_p]
        L.lh_model_info.argtypes = [C.c_void_p, _i, _i, C.POINTER(C.c_int64)]
        L.lh_trace.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(LhSource), C.c_int32,
                               C.c_uint64, C.POINTER(LhControls), C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        L.lh_trace_static.argtypes = L.lh_trace.argtypes
        L.lh_adjoint.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(LhSource),
                                 C.c_int32, C.c_uint64, C.POINTER(LhControls), C.c_void_p, C.c_int32, C.c_void_p,
                                 C.c_int32, C.c_void_p, C.c_void_p]
        L.lh_query.argtypes = [C.c_void_p, C.c_int32] + [C.c_void_p] * 4 + [C.c_float] + [C.c_void_p] * 6
        L.lh_batch_create.argtypes = [C.c_void_p, C.POINTER(LhSource), C.c_int32, C.c_void_p, C.c_int64, C.c_int32,
                                      C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]
        L.lh_batch_destroy.argtypes = [C.c_void_p]
        L.lh_batch_trace.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_uint64,
                                     C.POINTER(LhControls), C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        L.lh_batch_adjoint.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_uint64,
                                       C.POINTER(LhControls), C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                                       C.c_void_p, C.c_void_p]
        L.lh_last_kernel_ms.argtypes = [C.c_void_p, _f]
        L.lh_launches.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        L.lh_kernel_ms_of.argtypes = [C.c_void_p, C.c_int64, _f]
        L.lh_last_launch.argtypes = [C.c_void_p, _i]
        L.lh_model_jump.argtypes = [C.c_void_p, C.c_int32, _f, C.c_float, _i, _f]
        L.lh_bvh_build.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32] + [C.c_void_p] * 6 + [
            _i, _i]
        L.lh_last_error.restype = C.c_char_p
        for name in SYMBOLS:
            getattr(L, name).restype = C.c_int if name != "lh_last_error" else C.c_char_p
        _LIB = L
    return _LIB


def check(rc: int, what: str):
    """Raise EngineError for a non-zero return code."""
    if rc != 0:
        msg = lib().lh_last_error().decode()
        raise EngineError(f"{what}: error {rc}: {msg}")


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def geometry_struct(nav, sensor_of_vol, reach: float):
    """Return (LhGeometry, the arrays it points into) of NavArrays ``nav`` (the arrays must outlive the struct)."""
    h, T, K, G = nav.host, nav.csg["T"], nav.csg["K"], nav.csg["G"]
    f32 = lambda a: np.ascontiguousarray(np.asarray(a, np.float32))  # noqa: E731
    i32 = lambda a: np.ascontiguousarray(np.asarray(a, np.int32))  # noqa: E731
    n_leaves = len(T["typ"])
    keep = dict(
        R=f32(h["R"]).reshape(-1), t=f32(h["t"]).reshape(-1), kind=i32(h["kind"]), parent=i32(h["parent"]),
        cls=i32(h["cls"]), sensor=i32(sensor_of_vol), vg0=i32(h["vg0"]), vg1=i32(h["vg1"]),
        gR=f32(h["gR"]).reshape(-1), g0=i32(h["g0"]), g1=i32(h["g1"]), lo=f32(h["lo"]).reshape(-1),
        hi=f32(h["hi"]).reshape(-1), skip=i32(h["skip"]), start=i32(h["start"]), count=i32(h["count"]),
        items=i32(h["items"]), leaf_typ=i32(T["typ"]), leaf_p0=i32(T["p0"]), leaf_np=i32(T["np_"]),
        leaf_o=f32(np.stack([T["o0"], T["o1"], T["o2"]], 1)).reshape(-1),
        leaf_q=f32(np.stack([T["q0"], T["q1"], T["q2"]], 1)).reshape(-1),
        leaf_A=f32(np.stack([T[f"A{a}{b}"] for a in range(3) for b in range(3)], 1)).reshape(-1),
        leaf_b=f32(np.stack([T["b0"], T["b1"], T["b2"]], 1)).reshape(-1), leaf_c=f32(T["c"]),
        plane_n=f32(np.stack([T["n0"], T["n1"], T["n2"]], 1)).reshape(-1), plane_h=f32(T["h"]),
        kind_l0=i32(K["l0"]), kind_nl=i32(K["nl"]), kind_p0=i32(K["p0"]), kind_pl=i32(K["pl"]),
        prog_op=i32(G["op"]), prog_arg=i32(G["arg"]))
    if "skin" in h:  # the daughter-free cores (absent from the reference's arrays: no acceleration)
        keep["skin"] = i32(h["skin"])
    if "touch" in h:  # the volumes a sibling's box touches (absent: every exit looks for a sibling past the face)
        keep["touch"] = i32(h["touch"])
    lo_w, hi_w = getattr(nav, "lo_w", None), getattr(nav, "hi_w", None)
    if lo_w is not None and hi_w is not None and np.shape(lo_w) == (nav.n, 3) == np.shape(hi_w):
        keep["box_lo"], keep["box_hi"] = f32(lo_w).reshape(-1), f32(hi_w).reshape(-1)  # the reach bound's boxes
    n_nodes = len(keep["skip"]) if nav.stats.get("groups", 1) else 0
    g = LhGeometry(n_vol=nav.n, n_nodes=n_nodes, n_items=len(keep["items"]),
                   n_groups=int(nav.stats.get("groups", len(keep["g0"]))), n_leaves=n_leaves,
                   n_planes=len(keep["plane_h"]), n_prog=len(keep["prog_op"]), n_kinds=len(keep["kind_l0"]),
                   reach=float(reach))
    for k, v in keep.items():
        setattr(g, k, _ptr(v, C.c_float if v.dtype == np.float32 else C.c_int32))
    return g, keep
