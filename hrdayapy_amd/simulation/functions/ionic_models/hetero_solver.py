"""HeteroIonicSolver backed by one HIP kernel launch per step.

Drop-in for hrdayapy/simulation/functions/ionic_models/hetero_solver.py:57-211:
same constructor (Np, phi_myo, device, dtype, phi_endo_max, phi_epi_min, V0),
``init_states``, ``step(states, dt, I_stim) -> (I_ion, states)``,
``region_sizes`` and ``get_group_V``.  Instead of four index groups with a
gather / step / scatter each (hetero_solver.py:163-192), the whole node set
keeps one structure-of-arrays state buffer (18, N) and a per-node phenotype
byte, and a single kernel steps every node with its own phenotype.
"""
from __future__ import annotations

import numpy as np
import torch
from torch import Tensor

from .... import _native as nat
from ....ttp06 import INIT, N_STATES, STATE_NAMES

_PHI_ENDO_MAX = 0.35
_PHI_EPI_MIN = 0.65


class HeteroIonicSolver:
    def __init__(self, Np: int, phi_myo: np.ndarray, device, dtype=torch.float64,
                 phi_endo_max: float = _PHI_ENDO_MAX, phi_epi_min: float = _PHI_EPI_MIN,
                 V0: float = -86.709, verbose: bool = True):
        self.Np = int(Np)
        phi = np.asarray(phi_myo).astype(np.float32)
        self.N_myo = len(phi)
        self.N = self.Np + self.N_myo
        self.dev = torch.device(device)
        if self.dev.type != "cuda":
            raise RuntimeError("HeteroIonicSolver runs on the GPU only (HIP kernel)")
        self.dtype = dtype
        self.V0 = V0
        code = np.ones(self.N_myo, dtype=np.uint8)
        code[phi < phi_endo_max] = 0
        code[phi >= phi_epi_min] = 2
        ct = np.concatenate([np.zeros(self.Np, dtype=np.uint8), code])  # Purkinje -> endo
        self._ct = torch.from_numpy(ct).to(self.dev)
        self._groups = {"purkinje": np.arange(self.Np),
                        "endo": self.Np + np.nonzero(code == 0)[0],
                        "mid": self.Np + np.nonzero(code == 1)[0],
                        "epi": self.Np + np.nonzero(code == 2)[0]}
        st = torch.empty((N_STATES, self.N), dtype=dtype, device=self.dev)
        for k, name in enumerate(STATE_NAMES):
            st[k].fill_(INIT[name])
        self._state = st
        if verbose:
            tot = max(1, self.N)
            print(f"\n[HeteroIonicSolver] Transmural region assignment:")
            print(f"  phi thresholds: endo < {phi_endo_max:.2f} <= mid < {phi_epi_min:.2f} <= epi")
            for k, lab in (("purkinje", "Purkinje  (endo TTP06) "), ("endo", "Endo      (endo TTP06) "),
                           ("mid", "Mid-myo   (mcell TTP06)"), ("epi", "Epi       (epi TTP06)  ")):
                n = len(self._groups[k])
                print(f"  {lab}: {n:>8,}  ({100 * n / tot:.1f}%)")
            print(f"  Total nodes            : {self.N:>8,}\n")

    def init_states(self, N: int = None, dtype=None) -> dict[str, Tensor]:
        return {"V": torch.full((self.N,), self.V0, dtype=self.dtype, device=self.dev)}

    def step(self, states: dict[str, Tensor], dt: float, I_stim=None):
        V = states["V"]
        if V.dtype != self.dtype or V.shape[0] != self.N:
            raise ValueError(f"states['V'] must be ({self.N},) {self.dtype}")
        V = V.contiguous()
        I = torch.empty(self.N, dtype=self.dtype, device=self.dev)
        nat.call(f"hp_ttp06_step_{nat.sfx(self.dtype)}", V.data_ptr(), self._state.data_ptr(),
                 I.data_ptr(), self._ct.data_ptr(), 0, None, None, self.N, self.N, 1, float(dt),
                 0.0, 0.0, nat.stream_handle(self.dev))
        return I, states

    @property
    def region_sizes(self) -> dict[str, int]:
        return {k: len(v) for k, v in self._groups.items()}

    def get_group_V(self, group: str, V_global: Tensor) -> Tensor:
        idx = torch.as_tensor(self._groups[group], dtype=torch.long, device=V_global.device)
        return V_global[idx]

    def group_state(self, group: str) -> Tensor:
        """(N_group, 18) state of one region, in the reference's column order."""
        idx = torch.as_tensor(self._groups[group], dtype=torch.long, device=self.dev)
        return self._state[:, idx].t().contiguous()
