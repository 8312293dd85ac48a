"""Activation / deactivation times from Vm snapshots.

Drop-in for compute_activation_maps (hrdayapy/simulation/functions/
activation_maps.py:295-413): threshold crossings between consecutive frames,
linearly interpolated (:63-89), padded into (N_myo, n_events) arrays with
NaN, optional debounce.  The monodomain engine additionally records the
first activation on the device every step (MonodomainEngine.at_nodes).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np


def _interp_crossings(Vm, t, edge_mask, threshold):
    f, n = np.nonzero(edge_mask)
    if f.size == 0:
        return n, np.zeros(0, dtype=np.float32)
    V0, V1 = Vm[f, n], Vm[f + 1, n]
    den = V1 - V0
    den = np.where(den == 0, 1e-9, den)
    frac = np.clip((threshold - V0) / den, 0.0, 1.0)
    return n, (t[f] + frac * (t[f + 1] - t[f])).astype(np.float32)


def _pad_events(node, tc, N):
    order = np.lexsort((tc, node))
    node, tc = node[order], tc[order]
    counts = np.bincount(node, minlength=N)
    m = int(counts.max(initial=0))
    out = np.full((N, m), np.nan, dtype=np.float32)
    if node.size:
        start = np.concatenate([[0], np.cumsum(counts)[:-1]])
        col = np.arange(node.size) - start[node]
        out[node, col] = tc
    return out, counts.astype(np.int32)


def _apply_min_gap(times, gap):
    if gap is None or times.shape[1] < 2:
        return times
    out = times.copy()
    for i in range(out.shape[0]):
        kept, last = [], -np.inf
        for v in out[i]:
            if np.isfinite(v) and v - last >= gap:
                kept.append(v)
                last = v
        out[i] = np.nan
        out[i, :len(kept)] = kept
    return out


@dataclass
class ActivationMaps:
    activation_times: np.ndarray
    deactivation_times: np.ndarray
    n_activations: np.ndarray
    n_deactivations: np.ndarray
    coords_mm: np.ndarray
    vox_idx: np.ndarray
    act_threshold: float
    deact_threshold: float
    grid_shape: tuple = field(default=None)

    @property
    def N_myo(self):
        return self.activation_times.shape[0]

    def first_activation(self):
        a = self.activation_times
        return a[:, 0] if a.shape[1] else np.full(a.shape[0], np.nan, np.float32)


def compute_activation_maps(vm_snapshots, S=None, *, act_threshold=-40.0, deact_threshold=-40.0,
                            min_event_gap_ms=None, save_path=None) -> ActivationMaps:
    """vm_snapshots: a dict / npz with Vm (n_frames, N_myo), time_vm, coords_mm, vox_idx."""
    Vm = np.asarray(vm_snapshots["Vm"])
    t = np.asarray(vm_snapshots["time_vm"])
    if Vm.shape[0] < 2:
        raise ValueError("vm_snapshots has fewer than 2 frames -- cannot detect threshold crossings.")
    N = Vm.shape[1]
    above = Vm >= act_threshold
    na, ta = _interp_crossings(Vm, t, np.diff(above.astype(np.int8), axis=0) == 1, act_threshold)
    act, _ = _pad_events(na, ta, N)
    act = _apply_min_gap(act, min_event_gap_ms)
    above_d = Vm >= deact_threshold
    nd, td = _interp_crossings(Vm, t, np.diff(above_d.astype(np.int8), axis=0) == -1,
                               deact_threshold)
    deact, _ = _pad_events(nd, td, N)
    deact = _apply_min_gap(deact, min_event_gap_ms)
    vox = np.asarray(vm_snapshots["vox_idx"])
    shape = tuple(S.shape) if S is not None else tuple((vox.max(axis=0) + 1).tolist())
    maps = ActivationMaps(act, deact, (~np.isnan(act)).sum(axis=1).astype(np.int32),
                          (~np.isnan(deact)).sum(axis=1).astype(np.int32),
                          np.asarray(vm_snapshots["coords_mm"]), vox, act_threshold,
                          deact_threshold, shape)
    if save_path is not None:
        np.savez_compressed(Path(save_path).with_suffix(".npz"), activation_times=act,
                            deactivation_times=deact, n_activations=maps.n_activations,
                            n_deactivations=maps.n_deactivations, coords_mm=maps.coords_mm,
                            vox_idx=vox, grid_shape=np.array(shape, np.int32))
    return maps


def load_vm_snapshots(path):
    """Load vm_snapshots.npz (schema of purkinje_myocardium_anisotropic_solver.py:1975-1984)."""
    p = Path(path).with_suffix(".npz")
    if not p.exists():
        raise FileNotFoundError(f"No Vm snapshots at {p}")
    z = np.load(p)
    return {k: z[k] for k in z.files}
