from .gpu_ttp06 import EndoTTP06, EpiTTP06, MidTTP06, ttp06_step  # noqa: F401
from .hetero_solver import HeteroIonicSolver  # noqa: F401
from .protocol import RegionIonicModel  # noqa: F401
from .ttp06_phenotype_constants import PHENOTYPE_LABELS, TTP06_PHENOTYPES  # noqa: F401
