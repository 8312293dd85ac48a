"""voxel_size="from_geometry" resolution (hrdayapy/_spacing_resolve.py:24-73)."""
from __future__ import annotations

import numpy as np


def resolve_voxel_size(value, anatomy=None, param_name="voxel_size"):
    if isinstance(value, str):
        if value != "from_geometry":
            raise ValueError(f"{param_name} must be a float or 'from_geometry', got {value!r}")
        if anatomy is None or "spacing_mm" not in anatomy:
            raise ValueError(f"{param_name}='from_geometry' needs anatomy=... with 'spacing_mm'")
        sp = np.atleast_1d(np.asarray(anatomy["spacing_mm"], dtype=float))
        if not np.allclose(sp, sp[0], rtol=1e-6):
            raise ValueError(f"anisotropic spacing {sp.tolist()} cannot define a single {param_name}")
        return float(sp[0])
    return float(value)
