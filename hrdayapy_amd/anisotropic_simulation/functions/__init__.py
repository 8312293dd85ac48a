from .purkinje_myocardium_anisotropic_solver import (  # noqa: F401
    build_coupled_system, purkinje_myocardium_anisotropic_solver)
from ...simulation.functions.activation_maps import (  # noqa: F401
    ActivationMaps, VmSnapshots, compute_activation_maps, load_activation_maps,
    load_vm_snapshots, save_activation_maps)
