"""NodeIonicDispatcher: per-node model assignment with one HIP launch for the
TTP06 phenotypes.

Drop-in for hrdayapy/simulation/functions/ionic_models/dispatcher.py:69-209:
the same constructor (region_map, model_registry, device, dtype), ``step(V,
dt, I_stim) -> I_ion``, ``get_state``, ``set_state``, ``region_labels``,
``model_for_region``, ``node_indices_for_region`` and ``region_summary``.
Labels without a registry entry run SilentCell (zero current, no state).

Where the reference gathers, steps and scatters every region group in turn
(:140-178), here every node whose model is one of the HIP-backed TTP06
phenotypes (EndoTTP06 / MidTTP06 / EpiTTP06 of this package) is stepped by a
single hp_ttp06_step launch over a structure-of-arrays state buffer in node
order with a per-node phenotype byte; V is read in place when those nodes are
all the nodes.  Any other model (a user's RegionIonicModel) keeps the
reference's gather / step / scatter on its own group, and SilentCell nodes
launch nothing.  ``get_state(label)`` returns the (N_group, n_states) state of
a region in the group's node order, as the reference stores it.
"""
from __future__ import annotations

import numpy as np
import torch
from torch import Tensor

from .... import _native as nat
from ....ttp06 import INIT, N_STATES, STATE_NAMES
from .gpu_ttp06 import _CODE, EndoTTP06, EpiTTP06, MidTTP06
from .silent_cell import SilentCell

_NATIVE = (EndoTTP06, MidTTP06, EpiTTP06)


class NodeIonicDispatcher:
    def __init__(self, region_map: np.ndarray, model_registry: dict, device,
                 dtype: torch.dtype = torch.float32):
        rm = np.asarray(region_map)
        self.N = len(rm)
        self.device = torch.device(device)
        self.dtype = dtype
        silent = SilentCell()
        self._groups = []  # (label, model, idx tensor, native)
        native_nodes, native_code = [], []
        lines = []
        for label in np.unique(rm):
            model = model_registry.get(int(label), silent)
            where = np.nonzero(rm == label)[0]
            native = isinstance(model, _NATIVE)
            if native:
                native_nodes.append(where)
                native_code.append(np.full(len(where), _CODE[model._key], np.uint8))
            self._groups.append((int(label), model,
                                 torch.as_tensor(where, dtype=torch.long, device=self.device),
                                 native))
            lines.append(f"  region {int(label):>3d} : {model.name:<30s}  n={len(where):>8,}")
        self.region_summary = "\n".join(lines)
        # the native nodes in node order, with their phenotype codes
        nn = np.concatenate(native_nodes) if native_nodes else np.zeros(0, np.int64)
        cc = np.concatenate(native_code) if native_code else np.zeros(0, np.uint8)
        order = np.argsort(nn, kind="stable")
        nn, cc = nn[order], cc[order]
        self._n_native = len(nn)
        self._all_native = self._n_native == self.N
        if self._n_native and self.device.type != "cuda":
            raise RuntimeError("the TTP06 phenotypes run on the GPU only (HIP kernel)")
        self._native_idx = torch.as_tensor(nn, dtype=torch.long, device=self.device)
        self._native_code = torch.as_tensor(cc, dtype=torch.uint8, device=self.device)
        # position of every node in the native buffer (-1: not native)
        pos = np.full(self.N, -1, dtype=np.int64)
        pos[nn] = np.arange(len(nn))
        self._pos = pos
        self._soa = torch.empty((N_STATES, max(self._n_native, 1)), dtype=dtype,
                                device=self.device)
        for k, name in enumerate(STATE_NAMES):
            self._soa[k].fill_(INIT[name])
        self._state = {}
        for label, model, idx, native in self._groups:
            if not native:
                self._state[label] = model.init_state_tensor(idx.numel(), self.device, dtype)
        print("[NodeIonicDispatcher] Region -> model assignment:")
        print(self.region_summary)

    # ----------------------------------------------------------------- step
    def step(self, V: Tensor, dt: float, I_stim=0.0) -> Tensor:
        if V.shape[0] != self.N or V.dtype != self.dtype:
            raise ValueError(f"V must be ({self.N},) {self.dtype}")
        I_ion = torch.zeros(self.N, dtype=self.dtype, device=V.device)
        if self._n_native:
            Vn = V.contiguous() if self._all_native else V[self._native_idx].contiguous()
            In = I_ion if self._all_native else torch.empty(self._n_native, dtype=self.dtype,
                                                             device=V.device)
            nat.call(f"hp_ttp06_step_{nat.sfx(self.dtype)}", Vn.data_ptr(), self._soa.data_ptr(),
                     In.data_ptr(), self._native_code.data_ptr(), 0, None, None, self._n_native,
                     self._soa.shape[1], 1, float(dt), 0.0, 0.0, nat.stream_handle(V.device))
            if not self._all_native:
                I_ion[self._native_idx] = In
        for label, model, idx, native in self._groups:
            if native or isinstance(model, SilentCell) or idx.numel() == 0:
                continue
            stim = I_stim[idx] if isinstance(I_stim, Tensor) else I_stim
            I_g, st = model.step(V[idx], self._state[label], dt, stim)
            I_ion[idx] = I_g
            self._state[label] = st
        return I_ion

    # ---------------------------------------------------------------- state
    def _cols(self, label):
        for lab, model, idx, native in self._groups:
            if lab == label:
                return model, idx, native
        raise KeyError(f"Region label {label} not found in dispatcher.")

    def get_state(self, region_label: int) -> Tensor:
        """(N_group, n_states) state of a region (a copy for the TTP06 regions)."""
        model, idx, native = self._cols(region_label)
        if not native:
            return self._state[region_label]
        cols = torch.as_tensor(self._pos[idx.cpu().numpy()], dtype=torch.long, device=self.device)
        return self._soa[:, cols].t().contiguous()

    def set_state(self, region_label: int, state: Tensor) -> None:
        model, idx, native = self._cols(region_label)
        if not native:
            self._state[region_label] = state
            return
        if state.shape != (idx.numel(), N_STATES):
            raise ValueError(f"state must be ({idx.numel()}, {N_STATES})")
        cols = torch.as_tensor(self._pos[idx.cpu().numpy()], dtype=torch.long, device=self.device)
        self._soa[:, cols] = state.to(self.device, self.dtype).t()

    def region_labels(self) -> list[int]:
        return [lab for lab, _, _, _ in self._groups]

    def model_for_region(self, region_label: int):
        return self._cols(region_label)[0]

    def node_indices_for_region(self, region_label: int) -> Tensor:
        return self._cols(region_label)[1]
