from .activation_maps import (ActivationMaps, VmSnapshots, compute_activation_maps,  # noqa: F401
                              load_activation_maps, load_vm_snapshots, save_activation_maps)
