"""SilentCell: the inexcitable model (zero ionic current, no state).

Drop-in for hrdayapy/simulation/functions/ionic_models/silent_cell.py: same
``name``, ``state_names() -> []``, ``init_state_tensor(N, device, dtype) ->
(N, 0)`` and ``step(V, state, dt, I_stim) -> (zeros, state)``.  The
dispatcher never launches anything for its nodes: their current stays zero.
"""
from __future__ import annotations

import torch
from torch import Tensor


class SilentCell:
    name = "SilentCell"

    def state_names(self) -> list[str]:
        return []

    def init_state_tensor(self, N: int, device, dtype) -> Tensor:
        return torch.empty((N, 0), dtype=dtype, device=device)

    def step(self, V: Tensor, state: Tensor, dt: float, I_stim=None) -> tuple[Tensor, Tensor]:
        return torch.zeros(V.shape[0], dtype=V.dtype, device=V.device), state
