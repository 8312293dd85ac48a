from .generate_fibre_direction import generate_fibre_direction

__all__ = ["generate_fibre_direction"]
