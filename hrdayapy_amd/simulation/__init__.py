from .coupled import compute_coupled, load_coupled  # noqa: F401
from .functions.activation_maps import (ActivationMaps, VmSnapshots,  # noqa: F401
                                        compute_activation_maps, load_activation_maps,
                                        load_vm_snapshots, save_activation_maps)
