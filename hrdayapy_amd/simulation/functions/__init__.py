from .activation_maps import compute_activation_maps  # noqa: F401
