"""The anisotropic package's copy of the ionic plugin package
(hrdayapy/anisotropic_simulation/functions/ionic_models, identical in the
reference to the simulation one): the same classes, re-exported."""
from ....simulation.functions.ionic_models import *  # noqa: F401,F403
from ....simulation.functions.ionic_models import (  # noqa: F401
    EndoTTP06, EpiTTP06, HeteroIonicSolver, MidTTP06, NodeIonicDispatcher, RegionIonicModel,
    SilentCell, make_transmural_region_map, region_map_from_segmentation, validate_region_map)
