"""Detectors: the navigator-array format, the GDML-subset reader and the synthetic tanks."""

from pathlib import Path

from .gdml import Group, Recipe, load
from .model import Detector
from .navarrays import NavArrays
from .lar import lar_box, voxel_points
from .tanks import muon_track, water_tank

DATA = Path(__file__).resolve().parents[1] / "data"

EXAMPLE_TANK = Recipe("example_tank", "TankLV", (Group("pmt", "PMT", values=str(DATA / "toy_pmt.json")),))


def example_tank() -> Detector:
    """Return the reference's tutorial detector (``examples/data/example_tank.gdml`` with its recipe)."""
    return load(str(DATA / "example_tank.gdml"), EXAMPLE_TANK)


__all__ = ["Detector", "NavArrays", "Recipe", "Group", "load", "example_tank", "water_tank", "muon_track",
           "lar_box", "voxel_points"]
