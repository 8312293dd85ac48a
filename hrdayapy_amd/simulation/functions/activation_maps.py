"""Activation / deactivation times from Vm snapshots, and the snapshot reader.

Drop-in for compute_activation_maps and its helpers
(hrdayapy/simulation/functions/activation_maps.py:63-486 and the
anisotropic copy hrdayapy/anisotropic_simulation/functions/
activation_maps.py:63-560, which adds the optional dV/dt gate), and for
load_vm_snapshots / VmSnapshots (hrdayapy/simulation/functions/
load_vm_snapshots.py:49-240).

Events are threshold crossings between consecutive frames, linearly
interpolated inside the frame interval; rising crossings are activations,
falling ones deactivations.  Per node they are packed, in time order, into
an (N_myo, n_events) table padded with NaN.  A debounce drops an event
closer than ``min_event_gap_ms`` to the previous kept one (the table then
shrinks to the widest remaining row), and the optional velocity gates drop
crossings whose |dV/dt| over the bracketing interval is below
``min_upstroke_vel`` / ``min_downstroke_vel``.

This is host post-processing of a recorded history.  The monodomain engine
records the first activation of every voxel on the device in each step
(MonodomainEngine.at_nodes), without any snapshots.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

_SNAP_KEYS = ("node_ids", "coords_mm", "vox_idx", "Vm", "time_vm", "dt_solver", "vm_save_dt")
_MAP_KEYS = ("node_ids", "coords_mm", "vox_idx", "activation_times", "deactivation_times",
             "n_activations", "n_deactivations", "act_threshold", "deact_threshold")


# --------------------------------------------------------------------------- snapshots
@dataclass
class VmSnapshots:
    """The arrays of a vm_snapshots.npz file: node_ids (N_myo,), coords_mm
    (N_myo,3), vox_idx (N_myo,3), Vm (n_frames,N_myo) mV, time_vm (n_frames,)
    ms, dt_solver and vm_save_dt in ms.  Item access (``snap["Vm"]``) reads
    the same fields."""
    node_ids: np.ndarray
    coords_mm: np.ndarray
    vox_idx: np.ndarray
    Vm: np.ndarray
    time_vm: np.ndarray
    dt_solver: float
    vm_save_dt: float
    path: Path = field(default_factory=lambda: Path("."))

    def __getitem__(self, key):
        if key not in _SNAP_KEYS:
            raise KeyError(key)
        return getattr(self, key)

    @property
    def n_frames(self) -> int:
        return int(self.Vm.shape[0])

    @property
    def N_myo(self) -> int:
        return int(self.Vm.shape[1])

    def __repr__(self) -> str:
        t = self.time_vm
        span = f"{t[0]:.2f}..{t[-1]:.2f}" if len(t) else "empty"
        return (f"VmSnapshots(N_myo={self.N_myo:,}, n_frames={self.n_frames}, t=[{span}] ms, "
                f"vm_save_dt={self.vm_save_dt:.4g} ms)")

    def vm_at_node(self, local_id: int) -> np.ndarray:
        """Vm history of one myocardial node (local index 0..N_myo-1)."""
        if not 0 <= local_id < self.N_myo:
            raise IndexError(f"local_id {local_id} out of range [0, {self.N_myo})")
        return self.Vm[:, local_id]

    def nearest_node(self, point_mm) -> int:
        d = np.linalg.norm(self.coords_mm - np.asarray(point_mm, np.float32)[None, :], axis=1)
        return int(d.argmin())

    def field_at_time(self, t_ms: float):
        """(Vm of the frame nearest t_ms, that frame's time)."""
        k = int(np.argmin(np.abs(self.time_vm - t_ms)))
        return self.Vm[k], float(self.time_vm[k])

    def nodes_in_sphere(self, centre_mm, radius_mm: float) -> np.ndarray:
        d = np.linalg.norm(self.coords_mm - np.asarray(centre_mm, np.float32)[None, :], axis=1)
        return np.nonzero(d <= radius_mm)[0].astype(np.int32)

    def mean_vm_in_sphere(self, centre_mm, radius_mm: float) -> np.ndarray:
        ids = self.nodes_in_sphere(centre_mm, radius_mm)
        if ids.size == 0:
            return np.zeros(self.n_frames, dtype=np.float32)
        return self.Vm[:, ids].mean(axis=1)


def load_vm_snapshots(path) -> VmSnapshots:
    """Read vm_snapshots.npz (schema of purkinje_myocardium_anisotropic_solver.py:1975-1984)."""
    p = Path(path)
    if not p.suffix:
        p = p.with_suffix(".npz")
    if not p.exists():
        raise FileNotFoundError(f"vm_snapshots file not found: {p}")
    z = np.load(str(p))
    missing = set(_SNAP_KEYS) - set(z.files)
    if missing:
        raise KeyError(f"vm_snapshots NPZ is missing keys: {sorted(missing)} (found {z.files})")
    return VmSnapshots(z["node_ids"], z["coords_mm"], z["vox_idx"], z["Vm"], z["time_vm"],
                       float(z["dt_solver"]), float(z["vm_save_dt"]), p)


# --------------------------------------------------------------------------- events
def _crossings(Vm, t, edge, threshold):
    """Interpolated crossing time and mean dV/dt over the bracketing interval for
    every True entry of edge (n_frames-1, N): (node, time, velocity)."""
    f, n = np.nonzero(edge)
    if f.size == 0:
        return n, np.zeros(0, np.float32), np.zeros(0, np.float32)
    v0, v1 = Vm[f, n].astype(np.float64), Vm[f + 1, n].astype(np.float64)
    t0, t1 = t[f].astype(np.float64), t[f + 1].astype(np.float64)
    dv = v1 - v0
    frac = np.clip((threshold - v0) / np.where(dv == 0, 1e-9, dv), 0.0, 1.0)
    span = t1 - t0
    vel = dv / np.where(span == 0, 1e-9, span)
    return n, (t0 + frac * span).astype(np.float32), vel.astype(np.float32)


def _gate(node, tc, vel, min_abs_vel):
    if min_abs_vel is None or node.size == 0:
        return node, tc
    keep = np.abs(vel) >= min_abs_vel
    return node[keep], tc[keep]


def _table(node, tc, N):
    """(N, max events) NaN-padded table of event times in time order per node."""
    if node.size == 0:
        return np.zeros((N, 0), np.float32)
    order = np.lexsort((tc, node))
    node, tc = node[order], tc[order]
    counts = np.bincount(node, minlength=N)
    col = np.arange(node.size) - np.repeat(np.cumsum(counts) - counts, counts)
    out = np.full((N, int(counts.max())), np.nan, np.float32)
    out[node, col] = tc
    return out


def _debounce(table, gap):
    """Drop events closer than gap to the previous kept event of the same node;
    the table narrows to the widest remaining row."""
    if gap is None or table.shape[1] <= 1:
        return table
    rows = []
    for r in table:
        kept, last = [], -np.inf
        for v in r[np.isfinite(r)]:
            if v - last >= gap:
                kept.append(v)
                last = v
        rows.append(kept)
    out = np.full((table.shape[0], max((len(k) for k in rows), default=0)), np.nan, np.float32)
    for i, k in enumerate(rows):
        out[i, :len(k)] = k
    return out


# --------------------------------------------------------------------------- maps
@dataclass
class ActivationMaps:
    node_ids: np.ndarray
    coords_mm: np.ndarray
    vox_idx: np.ndarray
    activation_times: np.ndarray
    deactivation_times: np.ndarray
    n_activations: np.ndarray
    n_deactivations: np.ndarray
    act_threshold: float
    deact_threshold: float
    grid_shape: np.ndarray | None = None
    min_upstroke_vel: float | None = None
    min_downstroke_vel: float | None = None
    source_vm_path: str = ""
    path: Path = field(default_factory=lambda: Path("."))

    @property
    def N_myo(self) -> int:
        return int(self.activation_times.shape[0])

    @property
    def n_beats_activated(self) -> int:
        return int(self.activation_times.shape[1])

    @property
    def n_beats_deactivated(self) -> int:
        return int(self.deactivation_times.shape[1])

    @property
    def n_events(self) -> int:
        return max(self.n_beats_activated, self.n_beats_deactivated)

    def __repr__(self) -> str:
        return (f"ActivationMaps(N_myo={self.N_myo:,}, activation_events={self.n_beats_activated}, "
                f"deactivation_events={self.n_beats_deactivated}, "
                f"never_activated={int((self.n_activations == 0).sum()):,})")

    @staticmethod
    def _column(table, event):
        k = event - 1
        if k < 0 or k >= table.shape[1]:
            return np.full(table.shape[0], np.nan, np.float32)
        return table[:, k].copy()

    def activation_map(self, event: int = 1) -> np.ndarray:
        """Activation time of the event-th beat (1-based) per node, NaN if none."""
        return self._column(self.activation_times, event)

    def deactivation_map(self, event: int = 1) -> np.ndarray:
        return self._column(self.deactivation_times, event)

    def duration_map(self, event: int = 1) -> np.ndarray:
        return self.deactivation_map(event) - self.activation_map(event)

    def apd_map(self, event: int = 1) -> np.ndarray:
        return self.duration_map(event)

    def first_activation(self) -> np.ndarray:
        return self.activation_map(1)

    def field_on_grid(self, values) -> np.ndarray:
        """Scatter a per-node array onto the voxel grid (NaN background)."""
        if self.grid_shape is None:
            raise ValueError("grid_shape is not set: pass S to compute_activation_maps()")
        g = np.full(tuple(int(x) for x in self.grid_shape), np.nan, np.float32)
        v = self.vox_idx
        g[v[:, 0], v[:, 1], v[:, 2]] = values
        return g

    def activation_grid(self, event: int = 1) -> np.ndarray:
        return self.field_on_grid(self.activation_map(event))

    def deactivation_grid(self, event: int = 1) -> np.ndarray:
        return self.field_on_grid(self.deactivation_map(event))

    def apd_grid(self, event: int = 1) -> np.ndarray:
        return self.field_on_grid(self.apd_map(event))


def _as_snapshots(v):
    """VmSnapshots from a VmSnapshots, a dict or an npz; Vm may be a GPU tensor."""
    if isinstance(v, VmSnapshots):
        return v
    Vm = v["Vm"]
    if not (hasattr(Vm, "is_cuda") and Vm.is_cuda):
        Vm = np.asarray(Vm)
    n = Vm.shape[1]

    def get(k, default):
        return v[k] if k in v else default

    return VmSnapshots(np.asarray(get("node_ids", np.arange(n, dtype=np.int32))),
                       np.asarray(v["coords_mm"]), np.asarray(v["vox_idx"]), Vm,
                       np.asarray(v["time_vm"]), float(get("dt_solver", np.nan)),
                       float(get("vm_save_dt", np.nan)))


def _device_table(Vm, t, threshold, rising, min_vel, min_gap):
    """The same event table computed by hp_crossing_events_f32 on the device:
    Vm (n_frames, N) FP32 torch tensor on the GPU, t (n_frames,) host times."""
    import torch
    from ... import _native as nat
    nf, N = Vm.shape
    dev = Vm.device
    td = torch.as_tensor(np.asarray(t, dtype=np.float64), device=dev)
    mv = -1.0 if min_vel is None else float(min_vel)
    mg = -1.0 if min_gap is None else float(min_gap)
    counts = torch.zeros(N, dtype=torch.int32, device=dev)
    s = nat.stream_handle(dev)
    nat.call("hp_crossing_events_f32", Vm.data_ptr(), td.data_ptr(), nf, N, float(threshold),
             int(rising), mv, mg, counts.data_ptr(), None, 0, s)
    m = int(counts.max()) if N else 0
    table = torch.empty((N, m), dtype=torch.float32, device=dev)
    if m and N:
        nat.call("hp_crossing_events_f32", Vm.data_ptr(), td.data_ptr(), nf, N, float(threshold),
                 int(rising), mv, mg, None, table.data_ptr(), m, s)
    return table.cpu().numpy()


def compute_activation_maps(vm_snapshots, S=None, *, act_threshold=-40.0, deact_threshold=-40.0,
                            min_event_gap_ms=None, min_upstroke_vel=None, min_downstroke_vel=None,
                            save_path=None, device=None) -> ActivationMaps:
    """Activation/deactivation tables of a VmSnapshots (or a dict / npz with its keys).

    With ``device`` (e.g. "cuda"), or when Vm is already a GPU tensor, the tables
    are computed on the device by hp_crossing_events_f32 (identical results for
    FP32 snapshots); only the tables travel to the host."""
    snap = _as_snapshots(vm_snapshots)
    Vm, t = snap.Vm, np.asarray(snap.time_vm)
    on_dev = hasattr(Vm, "is_cuda") and Vm.is_cuda
    if on_dev and device is None:
        device = Vm.device
    if Vm.shape[0] < 2:
        raise ValueError("vm_snapshots has fewer than 2 frames -- cannot detect threshold "
                         "crossings. Re-run the simulation with a smaller vm_save_dt / longer T.")
    N = Vm.shape[1]
    if device is not None:
        import torch
        Vd = Vm if on_dev else torch.as_tensor(np.asarray(Vm, dtype=np.float32), device=device)
        Vd = Vd.to(torch.float32).contiguous()
        act = _device_table(Vd, t, act_threshold, True, min_upstroke_vel, min_event_gap_ms)
        deact = _device_table(Vd, t, deact_threshold, False, min_downstroke_vel,
                              min_event_gap_ms)
    else:
        Vm = np.asarray(Vm)
        up = np.diff((Vm >= act_threshold).astype(np.int8), axis=0) == 1
        act = _table(*_gate(*_crossings(Vm, t, up, act_threshold), min_upstroke_vel), N)
        act = _debounce(act, min_event_gap_ms)
        down = np.diff((Vm >= deact_threshold).astype(np.int8), axis=0) == -1
        deact = _table(*_gate(*_crossings(Vm, t, down, deact_threshold), min_downstroke_vel), N)
        deact = _debounce(deact, min_event_gap_ms)
    vox = np.asarray(snap.vox_idx)
    if S is not None:
        S = np.asarray(S)
        shape = np.array(S.shape, np.int32)
        outside = int((S[vox[:, 0], vox[:, 1], vox[:, 2]] == 0).sum())
        if outside:
            print(f"[compute_activation_maps] WARNING: {outside:,} myocardial nodes fall outside "
                  "the provided S mask")
    else:
        shape = (vox.max(axis=0) + 1).astype(np.int32)
    maps = ActivationMaps(
        np.asarray(snap.node_ids).astype(np.int32), np.asarray(snap.coords_mm, np.float32),
        vox.astype(np.int32), act, deact, (~np.isnan(act)).sum(axis=1).astype(np.int32),
        (~np.isnan(deact)).sum(axis=1).astype(np.int32), float(act_threshold),
        float(deact_threshold), shape,
        None if min_upstroke_vel is None else float(min_upstroke_vel),
        None if min_downstroke_vel is None else float(min_downstroke_vel), str(snap.path))
    if save_path is not None:
        save_activation_maps(maps, save_path)
    return maps


def save_activation_maps(maps: ActivationMaps, path) -> None:
    """activation_maps.npz with the reference's keys (activation_maps.py:484-520)."""
    p = Path(path)
    if not p.suffix:
        p = p.with_suffix(".npz")
    p.parent.mkdir(parents=True, exist_ok=True)
    d = dict(node_ids=maps.node_ids.astype(np.int32), coords_mm=maps.coords_mm.astype(np.float32),
             vox_idx=maps.vox_idx.astype(np.int32),
             activation_times=maps.activation_times.astype(np.float32),
             deactivation_times=maps.deactivation_times.astype(np.float32),
             n_activations=maps.n_activations.astype(np.int32),
             n_deactivations=maps.n_deactivations.astype(np.int32),
             act_threshold=np.float32(maps.act_threshold),
             deact_threshold=np.float32(maps.deact_threshold),
             min_upstroke_vel=np.float32(np.nan if maps.min_upstroke_vel is None
                                         else maps.min_upstroke_vel),
             min_downstroke_vel=np.float32(np.nan if maps.min_downstroke_vel is None
                                           else maps.min_downstroke_vel),
             source_vm_path=str(maps.source_vm_path))
    if maps.grid_shape is not None:
        d["grid_shape"] = np.asarray(maps.grid_shape, np.int32)
    np.savez_compressed(p, **d)
    maps.path = p


def load_activation_maps(path) -> ActivationMaps:
    p = Path(path)
    if not p.suffix:
        p = p.with_suffix(".npz")
    if not p.exists():
        raise FileNotFoundError(f"activation_maps file not found: {p}")
    z = np.load(str(p), allow_pickle=False)
    missing = set(_MAP_KEYS) - set(z.files)
    if missing:
        raise KeyError(f"activation_maps NPZ is missing keys: {sorted(missing)}")

    def opt(k):
        if k not in z.files or np.isnan(z[k]):
            return None
        return float(z[k])

    return ActivationMaps(z["node_ids"], z["coords_mm"], z["vox_idx"], z["activation_times"],
                          z["deactivation_times"], z["n_activations"], z["n_deactivations"],
                          float(z["act_threshold"]), float(z["deact_threshold"]),
                          z["grid_shape"] if "grid_shape" in z.files else None,
                          opt("min_upstroke_vel"), opt("min_downstroke_vel"),
                          str(z["source_vm_path"]) if "source_vm_path" in z.files else "", p)
