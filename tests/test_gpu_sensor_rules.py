"""The sensor-crediting rules on the device (the reference's transport/step.py:544-547, :574-612): the sensor sphere's
born, entry and exit cases against their closed forms and the oracle's photons, and the born deposit's derivative in
the entry efficiency through the adjoint."""

import numpy as np
import pytest
import torch

import lumen_amd as L
from lumen_amd.autograd import simulate_diff, table_tensor
from oracle import oracle as O
from toys import sensor_volume_toy

pytestmark = pytest.mark.gpu
SRC = L.sources.Point((0.0, 0.0, 0.0), lam=420.0, n_photons=1.0, time=2.5)
WANT = {"born": 0.6, "entry": 0.0, "exit": 0.6 * (1 - 0.3 - 0.4), "exit_off": 0.0}


@pytest.mark.parametrize("rule", sorted(WANT))
def test_sensor_rules_on_device(gpu, rule):
    print(f"[sensor rules] {rule}", flush=True)
    det = sensor_volume_toy(rule, L_abs=2.0)
    n = 1 << 18
    g = L.simulate(det, SRC, samples=n, seed=5)
    a = float(g.arrivals[0])
    assert abs(a - WANT[rule]) < 1e-6, (a, WANT[rule])
    if a > 0:
        t = 2.5 if rule == "born" else 2.5 + 0.3 * 1.33 / 0.299792458
        assert abs(float(g.mean_arrival_time[0]) - t) < 1e-4
    tal, h = O.OracleRun(det).trace(SRC, 1 << 14, seed=5)
    assert abs(g.health.records / n - float(h[10]) / (1 << 14)) < 1e-9  # one record per photon, or none
    assert abs(g.health.balance()) < 1e-5 * g.health.start_weight


def test_born_deposit_adjoint(gpu):
    """d(arrivals)/d(eff) of photons born in a sensitive volume: exactly 1 per photon of unit weight."""
    det = sensor_volume_toy("born", keyed=True)
    key = "surfaces/sensor/EFFICIENCY"
    t = table_tensor(det.params(), key)
    light = simulate_diff(det, SRC, {key: t}, samples=1 << 16, seed=3)
    light.arrivals.sum().backward()
    assert abs(float(t.grad.sum()) - 1.0) < 1e-5


def test_key_read_by_two_tables_gets_one_gradient(gpu):
    """A key packed once per table that reads it (here the entry and exit rows' shared EFFICIENCY, the exit row never
    reached) has one gradient: the adjoint's slot, not one copy per table (born photons: d arrivals / d eff = 1)."""
    det = sensor_volume_toy("born", keyed=True)
    from lumen_amd.pack import pack

    key = "surfaces/sensor/EFFICIENCY"
    pk = pack(det, grad_keys=(key,))
    assert (pk.slots >= 0).sum() == 2 and pk.n_grad == 1  # two copies, one slot
    t = table_tensor(det.params(), key)
    simulate_diff(det, SRC, {key: t}, samples=1 << 14, seed=4).arrivals.sum().backward()
    assert abs(float(t.grad.sum()) - 1.0) < 1e-5
