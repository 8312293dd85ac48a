from .purkinje_myocardium_anisotropic_solver import (  # noqa: F401
    build_coupled_system, purkinje_myocardium_anisotropic_solver)
