from .coupled import compute_coupled, load_coupled  # noqa: F401
from .functions.activation_maps import compute_activation_maps, load_vm_snapshots  # noqa: F401
