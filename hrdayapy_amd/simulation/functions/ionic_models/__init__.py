from .dispatcher import NodeIonicDispatcher  # noqa: F401
from .gpu_ttp06 import EndoTTP06, EpiTTP06, MidTTP06, ttp06_step  # noqa: F401
from .hetero_solver import HeteroIonicSolver  # noqa: F401
from .protocol import RegionIonicModel  # noqa: F401
from .region_map import (LABEL_BORDER, LABEL_ENDO, LABEL_EPI, LABEL_INFARCT,  # noqa: F401
                         LABEL_MCELL, PHI_ENDO_MAX, PHI_EPI_MIN, make_transmural_region_map,
                         region_map_from_segmentation, validate_region_map)
from .silent_cell import SilentCell  # noqa: F401
from .ttp06_phenotype_constants import PHENOTYPE_LABELS, TTP06_PHENOTYPES  # noqa: F401
