from .bspm import compute_bspm, load_bspm  # noqa: F401
from .torso_grid import compute_torso_grid  # noqa: F401
