"""Transmural coordinate on the GPU (the part of hrdayapy.coordinates that
the synthetic geometry needs)."""
from .functions import generate_transmural_coordinate
from .laplace import LaplaceSolver

__all__ = ["generate_transmural_coordinate", "LaplaceSolver"]
