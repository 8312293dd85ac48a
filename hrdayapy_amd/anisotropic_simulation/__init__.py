from .coupled import compute_anisotropic_coupled, load_anisotropic_coupled  # noqa: F401
