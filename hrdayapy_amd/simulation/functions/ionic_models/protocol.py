"""The ionic plugin protocol (hrdayapy/simulation/functions/ionic_models/protocol.py:72-121).

Every ionic model exposes ``name``, ``state_names()``, ``init_state_tensor(N,
device, dtype) -> (N, n_states)`` and ``step(V, state, dt, I_stim) -> (I_ion,
state)``; the PDE owns V.  The HIP-backed TTP06 classes satisfy it.
"""
from __future__ import annotations

from typing import Protocol, runtime_checkable

import torch
from torch import Tensor


@runtime_checkable
class RegionIonicModel(Protocol):
    name: str

    def state_names(self) -> list[str]: ...

    def init_state_tensor(self, N: int, device: torch.device, dtype: torch.dtype) -> Tensor: ...

    def step(self, V: Tensor, state: Tensor, dt: float, I_stim) -> tuple[Tensor, Tensor]: ...
