"""Brick layout of the myocardium in HBM.

The voxel mask S (Nx,Ny,Nz) is cut into bricks of (bx,by,bz) voxels; only
bricks holding at least one myocardial voxel are stored, in Morton (Z-curve)
order of the brick grid: a run of consecutive bricks is a compact block of
space, so the neighbour bricks whose faces a workgroup stages were mostly read
by nearby workgroups moments before (L2 reuse).  Every per-voxel array (V, the 18 TTP06 state variables, the
six conductivity-tensor components, the label byte) is a flat array of
n_bricks*bvol slots, slot = brick*bvol + (lx*by + ly)*bz + lz, so a wave of
64 lanes reads 64 consecutive slots of one brick (coalesced SoA).  A
27-entry neighbour table per brick lets a workgroup stage its halo tile.

Node order: node i is the i-th voxel of ``np.argwhere(S)`` -- the reference's
myocardial node numbering (purkinje_myocardium_anisotropic_solver.py:960) --
and ``slot_of_node[i]`` maps it to its slot.
"""
from __future__ import annotations

import numpy as np

BRICK_SHAPES = ((4, 4, 4), (8, 4, 4), (8, 8, 4), (8, 8, 8))
INACTIVE = 0xFF


def _spread3(v):
    """Bits of v (< 2^21) moved to every third position (Morton interleave)."""
    v = np.asarray(v, dtype=np.uint64) & np.uint64(0x1FFFFF)
    for sh, m in ((32, 0x1F00000000FFFF), (16, 0x1F0000FF0000FF), (8, 0x100F00F00F00F00F),
                  (4, 0x10C30C30C30C30C3), (2, 0x1249249249249249)):
        v = (v | (v << np.uint64(sh))) & np.uint64(m)
    return v


def morton3(x, y, z):
    """Morton code with z as the fastest-varying axis."""
    return (_spread3(x) << np.uint64(2)) | (_spread3(y) << np.uint64(1)) | _spread3(z)


class BrickLayout:
    def __init__(self, S: np.ndarray, brick=(8, 8, 8), order="morton"):
        brick = tuple(int(b) for b in brick)
        if brick not in BRICK_SHAPES:
            raise ValueError(f"brick shape {brick} not supported; choose one of {BRICK_SHAPES}")
        S = np.asarray(S).astype(bool)
        if S.ndim != 3:
            raise ValueError("S must be a 3-D voxel mask")
        self.shape = S.shape
        self.brick = brick
        bx, by, bz = brick
        self.bvol = bx * by * bz
        Nx, Ny, Nz = S.shape
        self.grid = (-(-Nx // bx), -(-Ny // by), -(-Nz // bz))
        gx, gy, gz = self.grid
        vox = np.argwhere(S).astype(np.int64)
        self.vox_idx = vox.astype(np.int32)
        self.n_nodes = len(vox)
        bc = vox // np.array(brick)
        bid = (bc[:, 0] * gy + bc[:, 1]) * gz + bc[:, 2]
        active = np.unique(bid)
        if order == "morton":
            a3 = np.stack([active // (gy * gz), (active // gz) % gy, active % gz], axis=1)
            active = active[np.argsort(morton3(a3[:, 0], a3[:, 1], a3[:, 2]), kind="stable")]
        elif order != "raster":
            raise ValueError(f"brick order {order!r}: 'morton' or 'raster'")
        self.order = order
        self.n_bricks = len(active)
        index_of = -np.ones(gx * gy * gz, dtype=np.int64)
        index_of[active] = np.arange(self.n_bricks)
        lc = vox % np.array(brick)
        local = (lc[:, 0] * by + lc[:, 1]) * bz + lc[:, 2]
        self.slot_of_node = (index_of[bid] * self.bvol + local).astype(np.int32)
        self.n_slots = self.n_bricks * self.bvol
        ab = np.stack([active // (gy * gz), (active // gz) % gy, active % gz], axis=1)
        self.brick_coords = ab.astype(np.int32)
        nbr = np.empty((self.n_bricks, 27), dtype=np.int32)
        k = 0
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    c = ab + np.array([dx, dy, dz])
                    ok = ((c[:, 0] >= 0) & (c[:, 0] < gx) & (c[:, 1] >= 0) & (c[:, 1] < gy)
                          & (c[:, 2] >= 0) & (c[:, 2] < gz))
                    lin = np.where(ok, (c[:, 0] * gy + c[:, 1]) * gz + c[:, 2], 0)
                    nbr[:, k] = np.where(ok, index_of[lin], -1)
                    k += 1
        self.nbr = nbr

    @classmethod
    def concat(cls, layouts):
        """Layouts of independent instances (same brick shape) back to back: bricks,
        slots and nodes of instance i follow those of instance i-1; neighbour
        entries keep pointing inside their own instance.  ``node_offset``,
        ``slot_offset`` and ``brick_offset`` give each instance's range."""
        layouts = list(layouts)
        if not layouts:
            raise ValueError("no layouts to concatenate")
        brick = layouts[0].brick
        if any(l.brick != brick for l in layouts):
            raise ValueError("all layouts of a concatenation need the same brick shape")
        out = cls.__new__(cls)
        out.shape = None
        out.brick = brick
        out.bvol = layouts[0].bvol
        out.grid = None
        nb = np.cumsum([0] + [l.n_bricks for l in layouts])
        nn = np.cumsum([0] + [l.n_nodes for l in layouts])
        out.brick_offset = nb
        out.slot_offset = nb * out.bvol
        out.node_offset = nn
        out.n_bricks = int(nb[-1])
        out.n_nodes = int(nn[-1])
        out.n_slots = out.n_bricks * out.bvol
        if out.n_slots > np.iinfo(np.int32).max:
            raise ValueError("more than 2^31 slots")
        out.vox_idx = np.concatenate([l.vox_idx for l in layouts])
        out.brick_coords = np.concatenate([l.brick_coords for l in layouts])
        out.slot_of_node = np.concatenate(
            [l.slot_of_node + np.int32(o) for l, o in zip(layouts, out.slot_offset[:-1])]
        ).astype(np.int32)
        out.nbr = np.concatenate(
            [np.where(l.nbr >= 0, l.nbr + np.int32(o), -1) for l, o in zip(layouts, nb[:-1])]
        ).astype(np.int32)
        return out

    @property
    def occupancy(self) -> float:
        return self.n_nodes / max(1, self.n_slots)

    def to_slots(self, values: np.ndarray, fill=0):
        """Scatter per-node values (N,) or (k,N) into slot arrays."""
        values = np.asarray(values)
        if values.ndim == 1:
            out = np.full(self.n_slots, fill, dtype=values.dtype)
            out[self.slot_of_node] = values
        else:
            out = np.full((values.shape[0], self.n_slots), fill, dtype=values.dtype)
            out[:, self.slot_of_node] = values
        return out

    def state_groups(self):
        """Compact cell-state index of the slots: (groups, state_of_node).
        groups[g] (int32 x4, one row per 64 consecutive slots) = (myocardial
        slots before slot 64*g, 0, myocardial mask bits 0-31, bits 32-63);
        state_of_node[i] = the state entry of node i = the rank of its slot
        among the myocardial slots."""
        if self.n_slots % 64:
            raise ValueError("the slot count must be a multiple of 64")
        tissue = np.zeros(self.n_slots, dtype=bool)
        tissue[self.slot_of_node] = True
        t = tissue.reshape(-1, 64)
        cnt = t.sum(axis=1)
        groups = np.zeros((len(t), 4), dtype=np.int64)
        groups[:, 0] = np.concatenate([[0], np.cumsum(cnt)[:-1]])
        w = (np.uint64(1) << np.arange(64, dtype=np.uint64))
        mask = (t.astype(np.uint64) * w).sum(axis=1, dtype=np.uint64)
        groups[:, 2] = (mask & np.uint64(0xFFFFFFFF)).astype(np.int64)
        groups[:, 3] = (mask >> np.uint64(32)).astype(np.int64)
        groups = groups.astype(np.uint32).view(np.int32).reshape(-1, 4)
        rank = np.cumsum(tissue) - 1
        return groups, rank[self.slot_of_node].astype(np.int32)

    def labels(self, celltype: np.ndarray, group: np.ndarray | None = None) -> np.ndarray:
        """Label byte per slot: phenotype in bits 0-1, stimulus group in bits 2-6."""
        lab = celltype.astype(np.uint8) & 3
        if group is not None:
            if group.max(initial=0) > 31:
                raise ValueError("at most 31 stimulus groups")
            lab = lab | (group.astype(np.uint8) << 2)
        return self.to_slots(lab, fill=INACTIVE).astype(np.uint8)


def choose_brick(S, target_blocks: int = 2048) -> tuple:
    """Default brick shape: the largest brick that still gives enough workgroups
    to fill the chip (>= target_blocks); otherwise the smallest brick.  S is the
    voxel mask or the voxel count."""
    n = int(S) if np.ndim(S) == 0 else int(np.count_nonzero(S))
    for b in reversed(BRICK_SHAPES):
        lay_blocks = n / (b[0] * b[1] * b[2])
        if lay_blocks >= target_blocks:
            return b
    return BRICK_SHAPES[0]
