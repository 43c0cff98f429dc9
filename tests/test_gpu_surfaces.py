"""The new surface types on the device (photocathode complex index, counting, transmitting mirrors, painted with
transmission) on the sensor sphere: against the closed forms and the oracle's same photons."""

import math

import numpy as np
import pytest

import lumen_amd as L
from oracle import oracle as O
from toys import specular_closed_form, sphere_toy

pytestmark = pytest.mark.gpu
A, L_ABS, N_IDX = 1.0, 10.0, 1.33
SRC = L.sources.Point((0.0, 0.0, 0.0), lam=420.0, n_photons=1.0)

CASES = {
    "photocathode": dict(surface="PHOTOCATHODE", R=0.0, eff=0.8, extra={"nr": 2.9, "nk": 1.6}),
    "counting": dict(surface="COUNTING", R=0.7, eff=0.6),
    "transmitting_mirror": dict(surface="SPECULAR_METAL", R=0.5, eff=0.9, extra={"T": 0.2}, transmitting=True,
                                far_opaque=True),
    "unified_normal_transmitting": dict(surface="UNIFIED_METAL_NORMAL", R=0.5, eff=0.9,
                                        extra={"T": 0.2, "p_spike": 0.3, "p_lobe": 0.3, "p_back": 0.4},
                                        transmitting=True, far_opaque=True),
    "painted_t": dict(surface="PAINTED_T", R=0.6, eff=1.0, extra={"T": 0.15}, far_opaque=True),
    "photocathode_per_normal": dict(surface="PHOTOCATHODE", R=0.0, eff=0.7, extra={"nr": 2.9, "nk": 1.6},
                                    per_normal=True),
}


def _closed(case):
    c = CASES[case]
    e = math.exp(-A / L_ABS)
    if case.startswith("photocathode"):
        nr, nk = c["extra"]["nr"], c["extra"]["nk"]
        rn = ((N_IDX - nr) ** 2 + nk**2) / ((N_IDX + nr) ** 2 + nk**2)
        if c.get("per_normal"):
            return c["eff"] * e / (1 - rn * math.exp(-2 * A / L_ABS))
        return c["eff"] * specular_closed_form(A, rn, L_ABS)
    if case == "counting":
        return c["eff"] * e
    R, T = c["R"], c["extra"]["T"]
    if case == "painted_t":
        b = 2 * A / L_ABS
        return c["eff"] * (1 - R - T) * e / (1 - R * 2 * (1 - (1 + b) * math.exp(-b)) / b**2)
    return c["eff"] * (1 - R - T) * e / (1 - R * math.exp(-2 * A / L_ABS))


@pytest.mark.parametrize("case", sorted(CASES))
def test_new_surface_types_on_device(gpu, case):
    print(f"[surfaces] {case}", flush=True)
    det = sphere_toy(A, L_abs=L_ABS, n=N_IDX, **CASES[case])
    n = 1 << 18
    g = L.simulate(det, SRC, samples=n, seed=4)
    q, err = float(g.arrivals[0]), float(g.arrivals_error[0])
    want = _closed(case)
    assert abs(q - want) < 2e-5 * want + 5 * err, (q, want, err)
    tal, h = O.OracleRun(det).trace(SRC, n, seed=4)
    assert abs(q - float(tal[0, 0])) < 2e-5 * want + 0.2 * err  # the same photons
    assert abs(g.health.balance()) < 1e-5 * g.health.start_weight
