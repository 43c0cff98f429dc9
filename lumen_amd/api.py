"""The drop-in surface: ``simulate(det, source, params, samples=, seed=) -> Light`` and ``Controls``.

Mirrors the reference's ``lumen.simulate`` (``api.py:876-1180``) for the path it runs on the HIP engine: the source's
photons are traced by the persistent kernel (``lh_trace``) into FP64 per-sensor tallies, from which ``Light`` gives
``arrivals``, ``arrivals_error`` (the per-photon second moment's error, ``readout/grid.py:48-60``),
``mean_arrival_time`` and the charge cells the readout reads, plus the run's ``health``.  Device memory and streams
are torch's; the results are torch tensors on the device (``Light.to_dlpack`` hands them on, JAX included).

Photon i of a run has the id ``first_id + i`` and its random numbers depend only on (seed, id, vertex), so a run's
photons are the same however it is split (the reference's ``first_id`` chunks, ``api.py:1045-1111``).
"""

from __future__ import annotations

import weakref
from dataclasses import dataclass, field
from typing import Any, NamedTuple

import numpy as np

from . import _lib, pack as packing
from .params import Params
from .sources import Batch

MAX_VERTICES = 1 << 22
ROULETTE = 0.01


class Controls(NamedTuple):
    """The step's controls (``transport/step.py:52-71``): ``max_distance`` (m; None: the geometry's reach),
    ``roulette`` (epsilon > 0), ``max_vertices`` (the guard), ``chunk`` (photons per work-queue claim; 0 the
    engine's default), ``blocks_per_cu`` (the persistent grid's resident blocks per CU; 0 as many as fit),
    ``tally_global`` (accumulate the tallies with global atomics even where a block's copy fits in LDS),
    ``nodes_global`` (read the BVH nodes from global memory even where they fit in the block's LDS), ``jump`` (the
    diffusion jump in Rayleigh-only media, ``transport/fastpath.py``: the reference's ``StepConfig(jump=True)``; its
    clearance grids are built on first use, ``lumen_amd.detector.clearance``), ``keep_zero_weight`` (trace a photon
    on after its live weight became exactly 0, as the reference does -- an edge row's defensive floor drew an outcome
    the live values give no probability; such a photon deposits nothing more, so only the ``records`` and
    ``vertices`` diagnostics change; off: it ends there), ``reach_bound`` (the reference's reach bound,
    ``geometry/reach.py``: a step whose free path ends inside its volume's box with no sensor volume's box ahead of
    it on the ray queries only a few bands past the free path; the same photons and results either way.  Off by
    default: on the water tank and the LAr box the query past the free path is cheap, and the bound's registers
    slow the loop by 14-20 %)."""  # noqa: D205, D209

    max_distance: Any = None
    roulette: float = ROULETTE
    max_vertices: int = MAX_VERTICES
    chunk: int = 0
    blocks_per_cu: int = 0
    tally_global: bool = False
    nodes_global: bool = False
    jump: bool = False
    keep_zero_weight: bool = False
    reach_bound: bool = False

    def struct(self) -> _lib.LhControls:
        if not (float(self.roulette) > 0.0):
            raise ValueError(f"Controls: the roulette is a number > 0, got {self.roulette!r}")
        return _lib.LhControls(float(self.roulette), 0.0 if self.max_distance is None else float(self.max_distance),
                               int(self.max_vertices), int(self.chunk), int(self.blocks_per_cu),
                               int(bool(self.tally_global)), int(bool(self.nodes_global)), int(bool(self.jump)),
                               int(bool(self.keep_zero_weight)), int(bool(self.reach_bound)))


class Health(NamedTuple):
    """A run's health (the reference's ``api.py:445-495``): every photon's start weight ends in one fate."""

    absorbed_bulk: float
    absorbed_surface: float
    lost: float
    escaped: float
    distance_cut: float
    abandoned: float
    truncated: float
    rouletted: float
    reweighted: float
    pushed: float
    records: float
    moment_spills: float
    vertices: float
    photons: float
    start_weight: float
    bounded: float = 0.0  # steps whose query the reach bound stopped a few bands past the free path

    FATES = ("absorbed_bulk", "absorbed_surface", "lost", "escaped", "distance_cut", "abandoned", "truncated",
             "rouletted", "reweighted")

    def balance(self) -> float:
        """Return the start weight minus every fate: 0 to float32 rounding per step."""
        return float(self.start_weight) - sum(float(getattr(self, k)) for k in self.FATES)

    @staticmethod
    def of(h: np.ndarray) -> "Health":
        return Health(*[float(x) for x in np.asarray(h, np.float64)[:16]])


@dataclass
class Light:
    """What ``simulate`` returns (the reference's ``api.py:515-537``): per sensor (with ``events``: per event and
    sensor) the expected arrivals, their Monte Carlo error, the mean arrival time; the charge cells (S, K) on the QE
    wavelength knots times the CE theta knots (``charge_axes`` = their counts, the wavelength's major); the raw
    tallies; the health."""  # noqa: D205, D209

    arrivals: Any
    arrivals_error: Any
    mean_arrival_time: Any
    charge: Any
    arrival2: Any
    health: Health
    samples: int
    provenance: dict = field(default_factory=dict)
    kernel_ms: float = float("nan")
    time: Any = None  # (S, nt) the time histogram's node masses (a timed detector: ``Detector.timed``)
    late: Any = None  # (S, 2) (q, q t) beyond the last node
    early: Any = None  # (S, 2) (q, q t) before the first node
    charge_axes: tuple = (1, 1)  # (wavelength knots, theta knots) of the charge cells
    sensor_group: Any = None  # a group's view (``groups``): its SensorGroup, whose keys ``readout.pe`` reads
    group_specs: list = field(default_factory=list)  # the detector's SensorGroups (``groups`` is built from them)
    n_events: int = 0  # E of a leading event axis (``simulate(events=E)`` or a Batch); 0: none

    @property
    def events(self) -> int:
        """The number of events of the outputs' leading event axis (0: none; the reference's ``Light.events``)."""
        return int(self.n_events)

    def event(self, e: int) -> "Light":
        """Return event ``e``'s Light (the reference's ``Light.event``, ``api.py:541-555``): every per-sensor array at
        that event; the health stays the run's."""  # noqa: D205, D209
        if not self.n_events:
            raise ValueError("Light.event: this light has no event axis (simulate(..., events=E) or a Batch)")
        if not 0 <= int(e) < self.n_events:
            raise IndexError(f"Light.event: event {e} of {self.n_events}")
        e = int(e)

        def cut(t):
            return None if t is None else t[e]

        return Light(cut(self.arrivals), cut(self.arrivals_error), cut(self.mean_arrival_time), cut(self.charge),
                     cut(self.arrival2), self.__dict__.get("_health"), self.samples,
                     provenance=dict(self.provenance, event=e), kernel_ms=self.__dict__.get("_kernel_ms"),
                     time=cut(self.time), late=cut(self.late), early=cut(self.early), charge_axes=self.charge_axes,
                     sensor_group=self.sensor_group, group_specs=self.group_specs)

    @property
    def groups(self) -> dict | None:
        """{group name: this Light restricted to the group's sensors} (the reference's ``Light.groups``,
        ``api.py:515-537``; None for a detector without groups).  Each view's sensor axis holds the group's sensors
        in the group's order and carries the group (``sensor_group``): the readout takes ``(light.group(name),
        params)``.  Built on first read and kept."""  # noqa: D205, D209
        if not self.group_specs:
            return None
        got = self.__dict__.get("_groups")
        if got is None:
            got = {g.name: self._view(g) for g in self.group_specs}
            self.__dict__["_groups"] = got
        return got

    def group(self, name: str) -> "Light":
        """Return the view of sensor group ``name`` (KeyError naming the groups there are)."""
        gs = self.groups
        if not gs or name not in gs:
            raise KeyError(f"Light: no sensor group {name!r}; the groups are {list(gs or ())}")
        return gs[name]

    def _view(self, g) -> "Light":
        """Return this Light on group g's sensors (gathered copies along the sensor axis)."""
        import torch

        s = int(self.arrivals.shape[-1])
        idx = np.asarray(g.sensors, np.int64)
        if idx.size and (idx.min() < 0 or idx.max() >= s):
            raise ValueError(f"Light: group {g.name!r} names sensors outside 0..{s - 1}")
        ix = torch.as_tensor(idx, device=self.arrivals.device)
        ev = self.arrivals.dim() - 1  # 1 with an event axis

        def cut(t):
            return None if t is None else t.index_select(ev, ix)

        return Light(cut(self.arrivals), cut(self.arrivals_error), cut(self.mean_arrival_time), cut(self.charge),
                     cut(self.arrival2), self.__dict__.get("_health"), self.samples,
                     provenance=dict(self.provenance, group=g.name), kernel_ms=self.__dict__.get("_kernel_ms"),
                     time=cut(self.time), late=cut(self.late), early=cut(self.early), charge_axes=self.charge_axes,
                     sensor_group=g, n_events=self.n_events)

    def to_dlpack(self, name="arrivals"):
        """Return a DLPack capsule of one output (device memory: ``jax.dlpack.from_dlpack`` takes it)."""
        import torch.utils.dlpack

        return torch.utils.dlpack.to_dlpack(getattr(self, name).contiguous())


def _lazy(name, convert):
    """A property over ``_<name>``: values that need the run to have finished (``simulate`` returns before it does)
    are converted on first read, by ``convert``, and kept."""  # noqa: D205, D209
    key = "_" + name

    def get(self):
        v = self.__dict__.get(key)
        if callable(v) and not isinstance(v, Health):
            v = convert(v())
            self.__dict__[key] = v
        return v

    def put(self, v):
        self.__dict__[key] = v

    return property(get, put)


# the run's health and kernel time: given as a callable by ``simulate`` (read after the kernel ends), else a value
Light.health = _lazy("health", lambda h: h if isinstance(h, Health) else Health.of(h))
Light.kernel_ms = _lazy("kernel_ms", float)


def moment_error(q2, s, n: int):
    """Return sqrt(n / (n - 1) (q2 - s^2 / n)) clipped at 0: the error of a sum of n i.i.d. photons from its per-photon
    second moment (``readout/grid.py:48-60``)."""  # noqa: D205, D209
    import torch

    n = float(max(int(n), 2))
    return torch.sqrt(torch.clamp(n / (n - 1.0) * (q2 - s * s / n), min=0.0))


class Engine:
    """A detector's model on the device (built once per detector and device) and its run state."""

    def __init__(self, det, device):
        self.det = det
        self.device = device
        layout = packing.pack(det).layout
        geo, self._keep = _lib.geometry_struct(det.nav, det.sensor_of_vol, det.reach)
        h = _lib.C.c_void_p()
        _lib.check(_lib.lib().lh_model_create(_lib.C.byref(geo), layout.ctypes.data_as(_lib._i), len(layout),
                                              _lib.C.byref(h)), "lh_model_create")
        self.handle = h
        self.layout = layout
        S, K, stride = _lib.C.c_int32(), _lib.C.c_int32(), _lib.C.c_int64()
        _lib.check(_lib.lib().lh_model_info(h, _lib.C.byref(S), _lib.C.byref(K), _lib.C.byref(stride)),
                   "lh_model_info")
        self.S, self.K, self.stride = S.value, K.value, stride.value
        self.Kt = int(layout[14])  # the charge cells: K = (QE wavelength knots) x (CE theta knots)
        self.Kl = self.K // self.Kt
        tg = getattr(det, "time_grid", None)
        self.nt = 0 if tg is None else tg.nt
        if self.stride != 3 * self.S + self.S * self.K + (self.S * (self.nt + 4) if self.nt else 0):
            raise _lib.EngineError(f"the engine's tally layout ({self.stride} cells per event) is not the detector's")
        weakref.finalize(self, _lib.lib().lh_model_destroy, h)
        self.jump_grids = None  # the diffusion jump's clearance grids, built on first use
        self._uploads = {}  # packed values by size: (the host array, its device copy), the last of each size

    def ensure_jump(self):
        """Build and upload the diffusion jump's clearance grids (once per engine); return them."""
        if self.jump_grids is None:
            from .detector.clearance import jump_grids

            self.jump_grids = jump_grids(self.det)
            for g in self.jump_grids:
                lo = (_lib.C.c_float * 3)(*[float(x) for x in g.lo])
                n = (_lib.C.c_int32 * 3)(*[int(x) for x in g.n])
                clear = np.ascontiguousarray(g.clear, np.float32)
                _lib.check(_lib.lib().lh_model_jump(self.handle, int(g.medium), lo, float(g.h), n,
                                                    clear.ctypes.data_as(_lib.C.POINTER(_lib.C.c_float))),
                           "lh_model_jump")
        return self.jump_grids

    def last_launch(self) -> dict:
        """Return the shape of the last launch on this model (grid, residency, LDS use, kernel features)."""
        out = np.zeros(8, np.int32)
        _lib.check(_lib.lib().lh_last_launch(self.handle, out.ctypes.data_as(_lib._i)), "lh_last_launch")
        keys = ("grid", "blocks_per_cu", "lds_bytes", "tally_in_lds", "nodes_in_lds", "lds_budget", "lds_per_cu",
                "features")
        return dict(zip(keys, (int(v) for v in out)))

    def query(self, x, d, vol=None, last=None, tmax=0.0):
        """Return (t, vol_after, normal, srf, located) of rays x, d (n, 3) on the device navigator (float32)."""
        import torch

        dev = self.device
        x = torch.as_tensor(np.asarray(x, np.float32).reshape(-1, 3), device=dev).contiguous()
        d = torch.as_tensor(np.asarray(d, np.float32).reshape(-1, 3), device=dev).contiguous()
        n = x.shape[0]
        vol = torch.full((n,), -1, dtype=torch.int32, device=dev) if vol is None else \
            torch.as_tensor(np.broadcast_to(np.asarray(vol, np.int32), (n,)).copy(), device=dev)
        last = torch.full((n,), -1, dtype=torch.int32, device=dev) if last is None else \
            torch.as_tensor(np.asarray(last, np.int32), device=dev)
        t = torch.zeros(n, dtype=torch.float32, device=dev)
        va = torch.zeros(n, dtype=torch.int32, device=dev)
        nrm = torch.zeros((n, 3), dtype=torch.float32, device=dev)
        srf = torch.zeros(n, dtype=torch.int32, device=dev)
        loc = torch.zeros(n, dtype=torch.int32, device=dev)
        st = torch.cuda.current_stream(dev)
        ptr = lambda a: _lib.C.c_void_p(a.data_ptr())  # noqa: E731
        _lib.check(_lib.lib().lh_query(self.handle, n, ptr(x), ptr(d), ptr(vol), ptr(last), float(tmax), ptr(t),
                                       ptr(va), ptr(nrm), ptr(srf), ptr(loc), _lib.C.c_void_p(st.cuda_stream)),
                   "lh_query")
        torch.cuda.synchronize(dev)
        return t.cpu().numpy(), va.cpu().numpy(), nrm.cpu().numpy(), srf.cpu().numpy(), loc.cpu().numpy()

    def kernel_ms(self, launch: int | None = None) -> float:
        """Return the kernel time (ms) of launch number ``launch`` on this model (None: the last); waits for it."""
        ms = _lib.C.c_float()
        if launch is None:
            _lib.check(_lib.lib().lh_last_kernel_ms(self.handle, _lib.C.byref(ms)), "lh_last_kernel_ms")
        else:
            _lib.check(_lib.lib().lh_kernel_ms_of(self.handle, int(launch), _lib.C.byref(ms)), "lh_kernel_ms_of")
        return ms.value

    def kernel_ms_sum(self, launches: range) -> float:
        """Return the summed kernel time (ms) of a range of launches; NaN once the oldest is past the model's last
        256 (its timing overwritten)."""  # noqa: D205, D209
        if self.launches() - launches.start > 256:
            return float("nan")
        return sum(self.kernel_ms(i) for i in launches)

    def launches(self) -> int:
        """Return the number of lh_trace / lh_adjoint launches made on this model."""
        n = _lib.C.c_int64()
        _lib.check(_lib.lib().lh_launches(self.handle, _lib.C.byref(n)), "lh_launches")
        return n.value

    def upload(self, values: np.ndarray):
        """Return the packed values on the device, the previous upload again when they are the same (a copy from
        pageable memory waits for the stream; unchanged values need none)."""  # noqa: D205, D209
        import torch

        last = self._uploads.get(values.nbytes)
        if last is not None and np.array_equal(last[0], values):
            return last[1]
        t = torch.as_tensor(values, device=self.device)
        self._uploads[values.nbytes] = (values.copy(), t)
        return t


_ENGINES: dict = {}


def engine(det, device="cuda"):
    """Return the detector's Engine on ``device`` (cached)."""
    import torch

    dev = torch.device(device)
    if dev.type != "cuda":
        raise RuntimeError("simulate: the photon loop runs on the GPU (HIP); there is no CPU path in the product")
    if not torch.cuda.is_available():
        raise RuntimeError("simulate: no GPU is visible (torch.cuda.is_available() is False)")
    key = (id(det), dev.index if dev.index is not None else torch.cuda.current_device())
    ent = _ENGINES.get(key)
    if ent is None or ent[0]() is not det:
        with torch.cuda.device(key[1]):
            eng = Engine(det, dev)
        _ENGINES[key] = (weakref.ref(det), eng)
        return eng
    return ent[1]


def source_rows(det, source, samples, first_id=0):
    """Return (the sources' rows as an LhSource array, the number of events)."""
    if isinstance(source, Batch):
        rows = (_lib.LhSource * len(source.sources))()
        first = int(first_id)
        n_ev = 0
        for i, (s, n) in enumerate(zip(source.sources, source.samples)):
            r = s.row(det, max(int(n), 1))
            r.n = int(n)
            r.first_id = first
            r.event = i if source.events is None else int(source.events[i])
            n_ev = max(n_ev, r.event + 1)
            first += int(n)
            rows[i] = r
        return rows, n_ev
    if samples is None or int(samples) < 1:
        raise ValueError("simulate: samples is a number of Monte Carlo photons >= 1")
    rows = (_lib.LhSource * 1)()
    r = source.row(det, int(samples))
    r.n = int(samples)
    r.first_id = int(first_id)
    rows[0] = r
    return rows, 1


def launch_rows(rows, chunk: int | None) -> list:
    """Return the launches of a run as LhSource arrays of at most ``chunk`` photons each (None: one launch of every
    row; the reference's ``readout/grid.py:358-362`` ``photon_chunks`` per source).  A row longer than a launch's room
    goes on in the next launch with its ``first_id`` advanced by the photons already taken, so every photon keeps its
    id and its random numbers; rows of no photons are dropped (a run of only such rows is one launch of them)."""  # noqa: D205, D209
    if chunk is None:
        return [rows]
    out, cur, room = [], [], chunk
    for r in rows:
        done, n = 0, int(r.n)
        while done < n:
            take = min(n - done, room)
            part = type(r).from_buffer_copy(r)
            part.n, part.first_id = take, int(r.first_id) + done
            cur.append(part)
            done, room = done + take, room - take
            if room == 0:
                out.append(cur)
                cur, room = [], chunk
    if cur:
        out.append(cur)
    if not out:
        return [rows]
    return [(type(rows[0]) * len(c))(*c) for c in out]


_GATED: dict = {}  # (id(detector), t0, dt, nt) -> (the detector, its timed copy): one model per gate


def gated(det, gate, step: float | None = None):
    """Return ``det`` with its time histogram over ``gate`` = (g0, g1) ns (the reference's ``simulate(gate=)``):
    nodes g0 + j dt up to the first at or past g1, dt ``step`` (default: the detector's own node spacing, else
    1 ns).  Light before or after the gate is in the early / late cells and in every total."""  # noqa: D205, D209
    import math

    g0, g1 = float(gate[0]), float(gate[1])
    if not (math.isfinite(g0) and math.isfinite(g1) and g1 > g0):
        raise ValueError(f"simulate: gate must be (g0, g1) with g1 > g0, got {gate!r}")
    dt = float(step) if step is not None else (det.time_grid.dt if det.time_grid is not None else 1.0)
    if not dt > 0:
        raise ValueError("simulate: the gate's node spacing must be > 0")
    nt = int(math.ceil((g1 - g0) / dt - 1e-9)) + 1
    k = (id(det), g0, dt, nt)
    hit = _GATED.get(k)
    if hit is None or hit[0] is not det:
        hit = _GATED[k] = (det, det.timed(g0, dt, nt))
    return hit[1]


def simulate(det, source, params: Params | None = None, *, samples: int | None = None, seed: int = 0,
             gate=None, events: int | None = None, reference: Params | None = None, controls: Controls | None = None,
             backend: str = "auto", chunk: int | None = None, records: int | None = None, diagnostics=(),
             truth=None, first_id: int = 0, edge_rows="structure", device="cuda", stream=None,
             gate_step: float | None = None) -> Light:
    """Trace ``samples`` Monte Carlo photons of ``source`` through ``det`` and return the light at each sensor.

    ``params``: the values (default ``det.params()``); ``reference``: the values every random choice is made at
    (default: ``params``) -- held fixed while ``params`` moves, the light at a seed is smooth in ``params`` (common
    random numbers).  A ``Batch`` source runs every source in one launch, one event each (or its ``events``
    labels): the outputs gain a leading event axis.  ``events`` (the reference's ``simulate(events=)``): E, the
    outputs' leading event axis for any source, every source's event below E.  ``gate`` = (g0, g1) ns: the time
    histogram's extent (``gated``; ``gate_step`` its node spacing).  ``edge_rows``: the surface rows drawn with the
    defensive mixture ("structure": the reference's rule; ``()``: none -- the same light in expectation, without the
    zero-weight photons an edge row sends on).

    ``chunk`` (the reference's memory lever): photons per launch; a run of more is several ``lh_trace`` launches
    added into the same tallies, each photon the same whatever the chunk (its id, hence its random numbers, is
    fixed), the results differing only by the order of the FP64 sums.  The engine's memory does not grow with the
    photons (no deposit records: every deposit is added into the tallies in place), so the default is one launch.
    ``records`` (the reference's deposit-record slots per launch): accepted (an int >= 1) and unused, for that
    reason; nothing can overflow.  ``backend``: "auto" (or "hip"), the HIP engine; the reference's "cpu" and "xla"
    photon loops have no counterpart here (the CPU oracle is a test checker, not a backend): refused.
    ``diagnostics`` and ``truth`` (the per-photon probes and the analog truth rows) are not built: refused unless
    empty.
    """
    import torch

    if backend not in ("auto", "hip"):
        raise ValueError(f"simulate: backend {backend!r}: this engine runs the photon loop on the GPU only "
                         f"('auto' or 'hip')")
    if diagnostics:
        raise ValueError("simulate: diagnostics (the per-photon probes) are not part of this engine")
    if truth is not None:
        raise ValueError("simulate: truth rows are not part of this engine")
    if chunk is not None and int(chunk) < 1:
        raise ValueError(f"simulate: chunk is a number of photons >= 1, got {chunk}")
    if records is not None and int(records) < 1:
        raise ValueError(f"simulate: records is a number of deposit records >= 1, got {records}")
    if gate is not None:
        det = gated(det, gate, gate_step)
    eng = engine(det, device)
    controls = Controls() if controls is None else controls
    p = det.params() if params is None else params
    pk = packing.pack(det, p, edge_rows=edge_rows)
    if not np.array_equal(pk.layout, eng.layout):
        raise ValueError("simulate: these values change the detector's structure (a table's knot count)")
    values = eng.upload(pk.values)
    ref_values = None
    if reference is not None:
        pr = packing.pack(det, reference, edge_rows=edge_rows)
        if not np.array_equal(pr.layout, eng.layout):
            raise ValueError("simulate: the reference's structure differs from the detector's")
        ref_values = torch.as_tensor(pr.values, device=eng.device)
    rows, n_ev = source_rows(det, source, samples, first_id)
    if events is not None:
        if int(events) < n_ev:
            raise ValueError(f"simulate: events={events} but a source's event is {n_ev - 1}")
        n_ev = int(events)
    # at least one cell: a detector without sensors has no tallies, and the C ABI takes no null buffer
    buf = torch.zeros((n_ev, max(eng.stride, 1)), dtype=torch.float64, device=eng.device)
    tally = buf[:, :eng.stride]
    health = torch.zeros(_lib.HEALTH_N, dtype=torch.float64, device=eng.device)
    current = torch.cuda.current_stream(eng.device)
    st = current if stream is None else stream
    if st != current:  # the buffers' fills and the values' uploads (on the current stream) before the kernel
        st.wait_stream(current)
    ctl = controls.struct()
    if controls.jump:
        eng.ensure_jump()
    first_launch = eng.launches()
    for part in launch_rows(rows, None if chunk is None else int(chunk)):
        _lib.check(_lib.lib().lh_trace(eng.handle, _lib.C.c_void_p(values.data_ptr()),
                                       None if ref_values is None else _lib.C.c_void_p(ref_values.data_ptr()),
                                       int(values.numel()), part, len(part),
                                       _lib.C.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF), _lib.C.byref(ctl),
                                       _lib.C.c_void_p(buf.data_ptr()), n_ev, _lib.C.c_void_p(health.data_ptr()),
                                       _lib.C.c_void_p(st.cuda_stream)), "lh_trace")
    launches = range(first_launch, eng.launches())
    if st != current:
        for t in (values, ref_values, buf, health):  # in use on `st` after this returns: kept from reuse meanwhile
            if t is not None:
                t.record_stream(st)
        current.wait_stream(st)  # the results, for what the caller does next on the current stream
    light = light_of(eng, tally, lambda: health.cpu().numpy(), samples if not isinstance(source, Batch) else None,
                     source, keep_events=events is not None)
    # the call returns before the kernels end: read on first use
    light.kernel_ms = lambda: eng.kernel_ms_sum(launches)
    light.group_specs = list(det.groups or [])
    tg = getattr(det, "time_grid", None)
    light.provenance = dict(seed=int(seed), samples=samples, detector=det.name, controls=controls._asdict(),
                            engine="liblumen_hip (gfx950)", reference=reference is not None, chunk=chunk,
                            launches=len(launches), time_grid=None if tg is None else (tg.t0, tg.dt, tg.nt))
    return light


def light_of(eng, tally, health, samples, source=None, keep_events=False) -> Light:
    """Return the Light of raw tallies (n_events, stride); one event of one source drops the event axis unless
    ``keep_events``."""  # noqa: D205, D209
    import torch

    S, K = eng.S, eng.K
    arr, arr_t, arr2 = tally[:, :S], tally[:, S:2 * S], tally[:, 2 * S:3 * S]
    charge = tally[:, 3 * S:3 * S + S * K].reshape(tally.shape[0], S, K)
    time = late = early = None
    if eng.nt > 0:
        t0 = 3 * S + S * K
        time = tally[:, t0:t0 + S * eng.nt].reshape(tally.shape[0], S, eng.nt)
        ov = tally[:, t0 + S * eng.nt:t0 + S * (eng.nt + 4)].reshape(tally.shape[0], S, 2, 2)
        late, early = ov[:, :, 0], ov[:, :, 1]
    if isinstance(source, Batch):
        # each event's photons are its sources' (a fixed count per source, not a draw from the mixture): the
        # moment error of an event uses the photons traced into it, the photon counts summed per event
        ev = np.arange(len(source.samples)) if source.events is None else np.asarray(source.events, np.int64)
        per_event = np.bincount(ev, weights=np.asarray(source.samples, np.float64), minlength=tally.shape[0])
        per_event = per_event[: tally.shape[0]]
        n = torch.as_tensor(np.maximum(per_event, 2.0), dtype=torch.float64, device=tally.device)[:, None]
        err = torch.sqrt(torch.clamp(n / (n - 1.0) * (arr2 - arr * arr / n), min=0.0))
    else:
        err = moment_error(arr2, arr, samples)
    mean_t = torch.where(arr > 0, arr_t / torch.where(arr > 0, arr, torch.ones_like(arr)), torch.zeros_like(arr))
    if tally.shape[0] == 1 and not isinstance(source, Batch) and not keep_events:
        arr, err, mean_t, charge, arr2 = arr[0], err[0], mean_t[0], charge[0], arr2[0]
        if time is not None:
            time, late, early = time[0], late[0], early[0]
    health = health if callable(health) else Health.of(health)
    n_events = 0 if arr.dim() == 1 else int(arr.shape[0])
    return Light(arr, err, mean_t, charge, arr2, health, samples or 0, time=time, late=late, early=early,
                 charge_axes=(eng.Kl, eng.Kt), n_events=n_events)
