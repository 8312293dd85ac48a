"""TTP06 phenotype classes backed by the HIP kernel (RegionIonicModel protocol).

Drop-in for hrdayapy/simulation/functions/ionic_models/gpu_ttp06.py:419-513:
same class names, ``name``, ``state_names()``, ``init_state_tensor(N, device,
dtype)`` and ``step(V, state, dt, I_stim) -> (I_ion, state)`` with state
(N, 18) in the reference's column order and I_ion in mV/ms.  ``step`` runs one
hp_ttp06_step kernel on the (N,18) tensor in place and returns it.  Tensors
must live on the GPU (float32 or float64); there is no CPU path.
"""
from __future__ import annotations

import torch
from torch import Tensor

from .... import _native as nat
from ....ttp06 import INIT, N_STATES, STATE_NAMES, TTP06_PHENOTYPES

_CODE = {"endo": 0, "mid": 1, "epi": 2}


def _init_tensor(N: int, device, dtype) -> Tensor:
    row = torch.tensor([INIT[k] for k in STATE_NAMES], dtype=dtype)
    return row.expand(N, N_STATES).contiguous().to(device)


def ttp06_step(V: Tensor, state: Tensor, dt: float, celltype, gks_scale=None, gto_scale=None):
    """One TTP06 step on the GPU.  ``celltype`` is an int phenotype code or a
    (N,) uint8 tensor of codes; ``state`` is (N,18) row-major, updated in place."""
    if not (V.is_cuda and state.is_cuda):
        raise RuntimeError("TTP06 step needs GPU tensors (the HIP kernel is the only implementation)")
    if state.dim() != 2 or state.shape[1] != N_STATES or state.shape[0] != V.shape[0]:
        raise ValueError(f"state must be (N, {N_STATES}) with N = len(V); got {tuple(state.shape)}")
    if state.dtype != V.dtype:
        raise ValueError("V and state must share a dtype")
    if not state.is_contiguous():
        raise ValueError("state must be contiguous (N,18)")
    V = V.contiguous()
    n = V.shape[0]
    I = torch.empty(n, dtype=V.dtype, device=V.device)
    ct = None
    fixed = 0
    if isinstance(celltype, Tensor):
        ct = celltype.to(device=V.device, dtype=torch.uint8).contiguous()
    else:
        fixed = int(celltype)
    for s in (gks_scale, gto_scale):
        if s is not None and (s.dtype != V.dtype or s.shape[0] != n):
            raise ValueError("per-node scales must match V in dtype and length")
    nat.call(f"hp_ttp06_step_{nat.sfx(V.dtype)}", V.data_ptr(), state.data_ptr(), I.data_ptr(),
             nat.ptr(ct), fixed, nat.ptr(gks_scale), nat.ptr(gto_scale), n, 1, N_STATES, float(dt),
             0.0, 0.0, nat.stream_handle(V.device))
    return I, state


class _Phenotype:
    _key = "endo"

    def state_names(self) -> list[str]:
        return list(STATE_NAMES)

    def init_state_tensor(self, N: int, device, dtype) -> Tensor:
        return _init_tensor(N, device, dtype)

    def step(self, V: Tensor, state: Tensor, dt: float, I_stim=None) -> tuple[Tensor, Tensor]:
        return ttp06_step(V, state, dt, _CODE[self._key])


class EndoTTP06(_Phenotype):
    """Endocardial TTP06 (ten Tusscher & Panfilov 2006, Table 1; endo s-gate)."""
    name = "TTP06_Endo"
    _key = "endo"
    _c = TTP06_PHENOTYPES["endo"]


class MidTTP06(_Phenotype):
    """Mid-myocardial (M-cell) TTP06 (Table 1; endo-type s-gate, large g_to)."""
    name = "TTP06_MCell"
    _key = "mid"
    _c = TTP06_PHENOTYPES["mid"]


class EpiTTP06(_Phenotype):
    """Epicardial TTP06 (Table 1; epi s-gate)."""
    name = "TTP06_Epi"
    _key = "epi"
    _c = TTP06_PHENOTYPES["epi"]
