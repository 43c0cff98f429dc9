"""lumen_amd: an MI355X-native photon-transport engine for the path behind ``lumen.simulate``.

    det = lumen_amd.load(path, recipe)          # or lumen_amd.detector.example_tank() / water_tank()
    light = lumen_amd.simulate(det, lumen_amd.sources.Point((0, 0, 0.1), lam=400.0, n_photons=2e5),
                               samples=2**18, seed=0)
    pe = lumen_amd.readout.pe(light, det.params(), det)

The photon loop is hand-written HIP for gfx950 (``lumen_amd/csrc``) behind the C ABI of ``include/lumen_hip.h``.
"""

from . import readout, sources
from .api import Controls, Health, Light, simulate
from .detector import Detector, Group, Recipe, example_tank, load, water_tank
from .params import Params, Table

__all__ = ["simulate", "Controls", "Light", "Health", "Params", "Table", "load", "Recipe", "Group", "Detector",
           "example_tank", "water_tank", "readout", "sources"]
