from .bspm import compute_bspm, load_bspm  # noqa: F401
from .registration import compute_registration, load_registration  # noqa: F401
from .torso_grid import compute_torso_grid, load_torso_grid  # noqa: F401
