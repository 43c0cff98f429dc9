"""Parameters: every value the photon loop reads, keyed by the detector's names (the reference's ``params.py:145-165``).

Keys: ``materials/<M>/<PROP>`` (a ``Table`` of photon energy (eV, ascending) and value, linear in energy between the
knots and held beyond them), ``surfaces/<S>/<PROP>`` (a Table or a number), ``sensors/<group>/<field>``
(``efficiency``: a number or one per sensor; ``qe``: a Table over wavelength in nm).  ``Params`` is immutable:
``set`` and ``scale`` return new Params, so a run's values are never changed behind it.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, NamedTuple

import numpy as np


class Table(NamedTuple):
    """A tabulated property: knots ``energy`` (eV, ascending; for a sensor's QE the wavelengths in nm) and ``value``."""

    energy: np.ndarray
    value: np.ndarray

    def scaled(self, s):
        """Return the table with every value times s."""
        return Table(np.asarray(self.energy, np.float64), np.asarray(self.value, np.float64) * s)


@dataclass(frozen=True)
class Info:
    """A key's unit, allowed range and gradient method (``Params.info``)."""

    unit: str
    range: tuple
    method: str


# a key's last part -> (unit, range, gradient method): the reference's catalogue for the keys this engine reads
_INFO = {
    "RINDEX": Info("-", (1.0, np.inf), "value only"),
    "ABSLENGTH": Info("m", (0.0, np.inf), "reweight"),
    "RAYLEIGH": Info("m", (0.0, np.inf), "reweight"),
    "REFLECTIVITY": Info("-", (0.0, 1.0), "reweight"),
    "TRANSMITTANCE": Info("-", (0.0, 1.0), "reweight"),
    "EFFICIENCY": Info("-", (0.0, 1.0), "reweight"),
    "SPECULARSPIKECONSTANT": Info("-", (0.0, 1.0), "reweight"),
    "SPECULARLOBECONSTANT": Info("-", (0.0, 1.0), "reweight"),
    "BACKSCATTERCONSTANT": Info("-", (0.0, 1.0), "reweight"),
    "sigma_alpha": Info("rad", (0.0, np.inf), "reweight"),
    "polish": Info("-", (0.0, 1.0 - 1e-3), "reweight"),
    "efficiency": Info("-", (0.0, 1.0), "after_tracking"),
    "qe": Info("-", (0.0, 1.0), "after_tracking"),
}


def _as_value(v):
    if isinstance(v, Table):
        return Table(np.asarray(v.energy, np.float64), np.asarray(v.value, np.float64))
    return np.asarray(v, np.float64)


class Params:
    """Parameter values by key (module docstring), with where each came from (``explain``)."""

    def __init__(self, values: dict, origin: dict | None = None):
        self._values = {k: _as_value(v) for k, v in values.items()}
        self._origin = dict(origin or {k: "detector file" for k in values})

    def keys(self):
        """Return the keys, sorted."""
        return sorted(self._values)

    def __contains__(self, key):
        return key in self._values

    def __getitem__(self, key) -> Any:
        try:
            return self._values[key]
        except KeyError:
            raise KeyError(f"Params: no key {key!r}; keys: {self.keys()}") from None

    def get(self, key, default=None):
        """Return the value of ``key`` or ``default``."""
        return self._values.get(key, default)

    def set(self, key, value) -> "Params":
        """Return Params with ``key`` set to ``value`` (a Table keeps its knots when given bare values)."""
        if key not in self._values:
            raise KeyError(f"Params.set: no key {key!r}")
        old = self._values[key]
        if isinstance(old, Table) and not isinstance(value, Table):
            value = Table(old.energy, np.broadcast_to(np.asarray(value, np.float64), old.value.shape).copy())
        vals = dict(self._values)
        vals[key] = _as_value(value)
        org = dict(self._origin)
        org[key] = "set"
        return Params(vals, org)

    def scale(self, key, s) -> "Params":
        """Return Params with ``key``'s values times s (every knot of a table)."""
        old = self[key]
        return self.set(key, old.scaled(s) if isinstance(old, Table) else old * s)

    def info(self, key) -> Info:
        """Return the key's unit, range and gradient method."""
        self[key]
        return _INFO.get(key.rsplit("/", 1)[-1], Info("-", (-np.inf, np.inf), "reweight"))

    def explain(self, key) -> str:
        """Return where the key's value came from."""
        self[key]
        return self._origin.get(key, "detector file")

    def flat(self) -> dict:
        """Return {key: float64 array of the values} (a Table's values only)."""
        return {k: (v.value if isinstance(v, Table) else np.atleast_1d(v)) for k, v in self._values.items()}

    def __repr__(self):
        return f"Params({len(self._values)} keys)"
