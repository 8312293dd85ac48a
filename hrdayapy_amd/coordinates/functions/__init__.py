from .generate_transmural_coordinate import generate_transmural_coordinate

__all__ = ["generate_transmural_coordinate"]
