"""Light sources: what ``simulate`` emits, as descriptors the kernel draws photons from on the device.

``Point`` (the reference's ``sources/point.py:38-86``: isotropic or a beam, monochromatic, here also a uniform
wavelength band) and ``CherenkovSegment`` (``sources/cherenkov.py:33-115``: Frank-Tamm light of a straight charged
segment with the medium's RINDEX: the wavelength from the proposal 1/lambda^2 over the RINDEX table's span, cos theta =
1/(beta n), the polarisation in the plane of the track and the photon, the start uniform along the segment).
``Batch`` is many sources in one launch, each with its own photon count, each its own tally event.

Photon i of a source uses only (seed, its id, vertex) random numbers; the direction law of an isotropic point is the
uniform sphere (the reference draws it from a normalised gaussian triple: the same law, other numbers).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib


@dataclass(frozen=True)
class Point:
    """Photons from ``position`` (m) at ``time`` (ns) standing for ``n_photons`` in expectation, at ``lam`` (nm) or
    uniform on ``lam_range`` (nm); ``direction`` None: isotropic, else a beam."""  # noqa: D205, D209

    position: tuple
    lam: float | None = None
    n_photons: float = 1.0
    time: float = 0.0
    direction: tuple | None = None
    lam_range: tuple | None = None

    def __post_init__(self):
        if (self.lam is None) == (self.lam_range is None):
            raise ValueError("Point: give a wavelength (lam) or a band (lam_range), not both")

    def row(self, det, samples: int) -> _lib.LhSource:
        s = _lib.LhSource()
        s.kind = _lib.LH_SRC_POINT if self.direction is None else _lib.LH_SRC_BEAM
        s.pos[:] = [float(v) for v in self.position]
        if self.direction is not None:
            d = np.asarray(self.direction, np.float64)
            s.dir[:] = list(d / np.linalg.norm(d))
        lo, hi = (self.lam, self.lam) if self.lam is not None else self.lam_range
        s.lam, s.lam_hi = float(lo), float(hi)
        s.time = float(self.time)
        s.weight = float(self.n_photons) / float(samples)
        return s


@dataclass(frozen=True)
class CherenkovSegment:
    """A charge ``charge`` (e) from ``start`` (m) along the unit ``direction`` for ``length`` (m) at ``beta`` in
    ``material``, at ``time`` (ns) at its start."""  # noqa: D205, D209

    material: str
    start: tuple = (0.0, 0.0, 0.0)
    direction: tuple = (0.0, 0.0, 1.0)
    length: float = 1.0
    beta: float = 1.0
    charge: float = 1.0
    time: float = 0.0

    def row(self, det, samples: int) -> _lib.LhSource:
        names = [m.name for m in det.media]
        if self.material not in names:
            raise ValueError(f"CherenkovSegment: no medium {self.material!r} in the detector ({names})")
        m = names.index(self.material)
        if det.media[m].rindex is None:
            raise ValueError(f"CherenkovSegment: {self.material!r} has no RINDEX")
        s = _lib.LhSource()
        s.kind = _lib.LH_SRC_CHERENKOV
        s.medium = m
        s.pos[:] = [float(v) for v in self.start]
        d = np.asarray(self.direction, np.float64)
        s.dir[:] = list(d / np.linalg.norm(d))
        s.length, s.beta, s.charge, s.time = float(self.length), float(self.beta), float(self.charge), float(self.time)
        s.weight = 1.0 / float(samples)
        return s


@dataclass(frozen=True)
class Batch:
    """Many sources in one launch: ``sources[i]`` traced with ``samples[i]`` photons into tally event i (or
    ``events[i]``)."""  # noqa: D205, D209

    sources: tuple
    samples: tuple
    events: tuple | None = field(default=None)

    def __post_init__(self):
        if len(self.sources) != len(self.samples) or len(self.sources) == 0:
            raise ValueError("Batch: one photon count per source, at least one source")
        if any(int(n) < 0 for n in self.samples):
            raise ValueError("Batch: photon counts are >= 0")
