"""Rule-based fibre field on the GPU (hrdayapy.fibres' generate_fibre_direction)."""
from .functions import generate_fibre_direction

__all__ = ["generate_fibre_direction"]
