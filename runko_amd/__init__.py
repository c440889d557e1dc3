"""runko_amd -- MI355X-native PIC time-step hot path with runko's Python API.

Drop-in use for runko drivers:  ``import runko_amd as runko``.

Exports the names a runko driver uses on the PIC time-step path:
``pic.threeD.Tile``, ``pic.threeD.ParticleState``, ``pic.threeD.ParticleStateBatch``,
``pic.threeD.reflector_wall``, ``emf.threeD.Tile``, ``emf.threeD.edge_bc``,
``tools.comm_mode``, ``Configuration``, ``TileGrid``, ``Simulation``,
``MovingInjector``, ``sample_boosted_juttner_synge``, ``runko_logger``,
``on_main_rank``.
The compute path is the HIP library librunko_amd.so (gfx950); it is required.
"""

import logging

from . import tools
from .configuration import Configuration
from .method_wrapper import MethodWrapper
from .moving_injector import MovingInjector
from .sampling import sample_boosted_juttner_synge
from .simulation import Simulation
from .tile_grid import TileGrid
from .tiles import EmfTile, ParticleState, ParticleStateBatch, PicTile, make_modules


class _Module:
    pass


pic = _Module()
emf = _Module()
pic.threeD, emf.threeD = make_modules()
comm_mode = tools.comm_mode


def on_main_rank():
    return True


def runko_logger(name="runko"):
    return logging.getLogger(name)


__all__ = ["pic", "emf", "tools", "comm_mode", "Configuration", "TileGrid", "Simulation", "MethodWrapper",
           "sample_boosted_juttner_synge", "on_main_rank", "runko_logger", "EmfTile", "PicTile",
           "ParticleState", "ParticleStateBatch", "MovingInjector"]
