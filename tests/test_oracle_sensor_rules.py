"""The sensor-crediting rules (the reference's transport/step.py:544-547 and :574-612, build/lookup.py:413-417) in
the oracle, on the sensor sphere in water (every photon from its centre, matched indices): a photon emitted inside a
sensitive volume counting by volume deposits the entry's detected share at its emission and ends; one counting only
entries leaves uncounted; a detecting border out of a sensor volume credits the volume left (eff (1 - R - T) per
photon, at the sphere's radius), and without the rule the water entered, which holds no sensor, gets nothing."""

import numpy as np
import pytest

import lumen_amd as L
from oracle import oracle as O
from toys import sensor_volume_toy

N_IDX = 1.33
SRC = L.sources.Point((0.0, 0.0, 0.0), lam=420.0, n_photons=1.0, time=2.5)


def _run(rule, n=1 << 14, **kw):
    det = sensor_volume_toy(rule, **kw)
    tal, h = O.OracleRun(det).trace(SRC, n, seed=5)
    return O.light_of(tal, h, 1, 1, n), h


def test_born_inside_a_sensitive_volume_is_counted_at_emission():
    o, h = _run("born")
    assert abs(o["arrivals"][0] - 0.6) < 1e-7  # eff as its float32 value
    assert abs(o["mean_arrival_time"][0] - 2.5) < 1e-9  # at the emission's time
    assert abs(h[1] - 1.0) < 1e-12 and h[10] == 1 << 14 and h[12] == 1 << 14  # absorbed there, one vertex each


def test_born_inside_a_volume_counting_entries_leaves_uncounted():
    o, h = _run("entry")
    assert o["arrivals"][0] == 0.0 and h[10] == 0


@pytest.mark.parametrize("L_abs", [None, 2.0])
def test_exit_sensor_credits_the_volume_left(L_abs):
    o, h = _run("exit", L_abs=L_abs)
    assert abs(o["arrivals"][0] - 0.6 * (1 - 0.3 - 0.4)) < 1e-7  # the sphere holds no absorber: no attenuation
    assert abs(o["mean_arrival_time"][0] - 2.5 - 0.3 * N_IDX / 0.299792458) < 1e-6  # at the border, flat index
    off, _ = _run("exit_off", L_abs=L_abs)
    assert off["arrivals"][0] == 0.0


GDML_SENSORS = """<?xml version="1.0" encoding="UTF-8"?>
<gdml>
<define>
  <matrix name="RIDX" coldim="2" values="1.0*eV 1.33 12.0*eV 1.33"/>
  <matrix name="R3" coldim="2" values="1.0*eV 0.3 12.0*eV 0.3"/>
  <matrix name="T4" coldim="2" values="1.0*eV 0.4 12.0*eV 0.4"/>
  <matrix name="E6" coldim="2" values="1.0*eV 0.6 12.0*eV 0.6"/>
</define>
<materials>
  <material name="Liq"><property name="RINDEX" ref="RIDX"/></material>
  <material name="Cell"><property name="RINDEX" ref="RIDX"/></material>
</materials>
<solids>
  <box name="WorldBox" x="4" y="4" z="4" lunit="m"/>
  <orb name="O" r="300" lunit="mm"/>
  <opticalsurface name="Out" model="glisur" finish="polished" type="dielectric_dielectric" value="1">
    <property name="REFLECTIVITY" ref="R3"/><property name="TRANSMITTANCE" ref="T4"/>
    <property name="EFFICIENCY" ref="E6"/>
  </opticalsurface>
</solids>
<structure>
  <volume name="cellV"><materialref ref="Cell"/><solidref ref="O"/></volume>
  <volume name="World"><materialref ref="Liq"/><solidref ref="WorldBox"/>
    <physvol name="cellA"><volumeref ref="cellV"/><position name="pa" x="-1" y="0" z="0" unit="m"/></physvol>
    <physvol name="cellB"><volumeref ref="cellV"/><position name="pb" x="1" y="0" z="0" unit="m"/></physvol>
  </volume>
  <bordersurface name="outA" surfaceproperty="Out"><physvolref ref="cellA"/><physvolref ref="World"/></bordersurface>
</structure>
<setup name="Default" version="1.0"><world ref="World"/></setup>
</gdml>
"""


@pytest.mark.parametrize("rule,count", [("any_entry", "volume"), ("any_entry", "entry"), ("border_into", "volume")])
def test_gdml_sensor_groups_build_the_crediting_tables(tmp_path, rule, count):
    """A recipe's sensor group: any_entry makes every entry a counting row (eff 1 without a value set) and, counting
    by volume, a born row; a detecting border out of a sensor placement (cellA -> World) is an exit sensor whatever
    the rule.  Traced from inside each cell: cellA's photons deposit at birth (any_entry by volume: 1 each) or at its
    border (eff (1 - R - T) = 0.18), cellB's own only at birth.  A class holds the placements of one signature:
    cellA's border out makes its class differ from cellB's."""
    from lumen_amd.detector.gdml import Group, Recipe, load

    f = tmp_path / "sensors.gdml"
    f.write_text(GDML_SENSORS)
    det = load(str(f), Recipe("cells", "World", (Group("cells", "cellV", rule=rule, count=count),)))
    assert det.credit_left is not None and det.credit_left.sum() == 1
    assert len({det.nav.host["cls"][v] for v in range(det.nav.n) if det.sensor_of_vol[v] >= 0}) == 2
    born = rule == "any_entry" and count == "volume"
    assert (det.born_row is not None) == born
    n = 1 << 12
    for x, sensor in ((-1.0, "cellA"), (1.0, "cellB")):
        src = L.sources.Point((x, 0.0, 0.0), lam=420.0, n_photons=1.0)
        tal, h = O.OracleRun(det).trace(src, n, seed=2)
        a = O.light_of(tal, h, 2, 1, n)["arrivals"]
        i = det.sensor_names.index(next(s for s in det.sensor_names if sensor in s))
        want = 1.0 if born else (0.6 * (1 - 0.3 - 0.4) if sensor == "cellA" else 0.0)
        assert abs(a[i] - want) < 1e-6, (sensor, a)  # the other cell sees the light that reaches it (cross-talk)
