"""CPU checks of the brick layout and of the C-ABI library (no GPU calls)."""
import re
from pathlib import Path

import numpy as np
import pytest

from hrdayapy_amd import _native as nat
from hrdayapy_amd.layout import BRICK_SHAPES, BrickLayout
from tests._problems import slab

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("order", ["morton", "raster"])
@pytest.mark.parametrize("brick", BRICK_SHAPES)
def test_layout_slots_and_neighbours(brick, order):
    S, _, _ = slab(13, 11, 7)
    S[4:7, 3:9, 2:5] = 0
    lay = BrickLayout(S, brick, order=order)
    vox = lay.vox_idx
    assert lay.n_nodes == S.sum()
    assert len(np.unique(lay.slot_of_node)) == lay.n_nodes
    bx, by, bz = brick
    # slot -> voxel coordinates through the brick table reproduce vox_idx
    b = lay.slot_of_node // lay.bvol
    loc = lay.slot_of_node % lay.bvol
    lx, ly, lz = loc // (by * bz), (loc // bz) % by, loc % bz
    xyz = lay.brick_coords[b] * np.array(brick) + np.stack([lx, ly, lz], 1)
    assert np.array_equal(xyz, vox)
    # neighbour table: centre entry is the brick itself, others are spatial neighbours
    assert np.array_equal(lay.nbr[:, 13], np.arange(lay.n_bricks))
    for k, (dx, dy, dz) in enumerate([(a, c, d) for a in (-1, 0, 1) for c in (-1, 0, 1)
                                      for d in (-1, 0, 1)]):
        ok = lay.nbr[:, k] >= 0
        assert np.array_equal(lay.brick_coords[lay.nbr[ok, k]],
                              lay.brick_coords[ok] + np.array([dx, dy, dz]))
    # absent entries are bricks without tissue
    stored = {tuple(c) for c in lay.brick_coords}
    for k, (dx, dy, dz) in enumerate([(a, c, d) for a in (-1, 0, 1) for c in (-1, 0, 1)
                                      for d in (-1, 0, 1)]):
        for i in np.nonzero(lay.nbr[:, k] < 0)[0]:
            assert tuple(lay.brick_coords[i] + np.array([dx, dy, dz])) not in stored
    # brick order: Morton codes (bits interleaved x, y, z from the top) ascending
    c = lay.brick_coords.astype(np.int64)
    key = np.zeros(len(c), np.int64)
    for bit in range(6):
        for ax in range(3):
            key |= ((c[:, ax] >> bit) & 1) << (3 * bit + 2 - ax)
    lin = (c[:, 0] * lay.grid[1] + c[:, 1]) * lay.grid[2] + c[:, 2]
    assert np.all(np.diff(key if order == "morton" else lin) > 0)
    lab = lay.labels(np.full(lay.n_nodes, 2, np.uint8), np.ones(lay.n_nodes, np.int64))
    assert (lab != 0xFF).sum() == lay.n_nodes
    assert np.all(lab[lay.slot_of_node] == (2 | (1 << 2)))


def test_library_exports_every_header_symbol():
    if not nat.lib_path().exists():
        pytest.skip("library not built")
    hdr = (ROOT / "include" / "hrdaya_hip.h").read_text()
    names = set(re.findall(r"^int (hp_\w+)\(", hdr, flags=re.M))
    assert names, "no declarations parsed"
    lib = nat.load()
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert names <= set(nat.exported_symbols())
    assert lib.hp_abi_version() == nat.ABI_VERSION


def test_concat_layout_keeps_instances_apart():
    """Concatenated instances: slots/nodes offset per instance, neighbour table
    entries stay inside their own instance, and each instance's part equals its
    own layout shifted by the offsets."""
    parts = [BrickLayout(slab(20, 16, 8)[0], (4, 4, 4)),
             BrickLayout(slab(9, 13, 5, seed=2)[0], (4, 4, 4)),
             BrickLayout(slab(4, 4, 4)[0], (4, 4, 4))]
    L = BrickLayout.concat(parts)
    assert L.n_nodes == sum(p.n_nodes for p in parts)
    assert L.n_slots == sum(p.n_slots for p in parts)
    for i, p in enumerate(parts):
        b0, b1 = L.brick_offset[i], L.brick_offset[i + 1]
        nb = L.nbr[b0:b1]
        assert np.array_equal(nb, np.where(p.nbr >= 0, p.nbr + b0, -1))
        n0, n1 = L.node_offset[i], L.node_offset[i + 1]
        assert np.array_equal(L.slot_of_node[n0:n1], p.slot_of_node + L.slot_offset[i])
    with pytest.raises(ValueError):
        BrickLayout.concat([parts[0], BrickLayout(slab(8, 8, 8)[0], (8, 8, 8))])


@pytest.mark.parametrize("brick", BRICK_SHAPES)
def test_state_groups_index_the_compact_state(brick):
    """state_groups: the compact state entry of each node is the rank of its
    slot among the myocardial slots (a permutation of 0..N-1 in slot order),
    and groups[s >> 6] gives it as base + popcount(mask below s & 63), with
    the mask bit set exactly on myocardial slots; concatenations included."""
    S, _, _ = slab(13, 11, 7)
    S[4:7, 3:9, 2:5] = 0
    for lay in (BrickLayout(S, brick),
                BrickLayout.concat([BrickLayout(S, brick), BrickLayout(slab(9, 5, 6)[0], brick)])):
        groups, son = lay.state_groups()
        assert groups.shape == (lay.n_slots // 64, 4) and groups.dtype == np.int32
        order = np.argsort(lay.slot_of_node)
        assert np.array_equal(son[order], np.arange(lay.n_nodes))
        g = groups.astype(np.int64) & 0xFFFFFFFF
        mask = g[:, 2] | (g[:, 3] << 32)
        tissue = np.zeros(lay.n_slots, bool)
        tissue[lay.slot_of_node] = True
        bits = (mask[:, None] >> np.arange(64)[None, :]) & 1
        assert np.array_equal(bits.ravel().astype(bool), tissue)
        s = lay.slot_of_node.astype(np.int64)
        below = mask[s >> 6] & ((np.int64(1) << (s & 63)) - 1)
        pc = np.array([bin(int(v) & (2**64 - 1)).count("1") for v in below])
        assert np.array_equal(groups[s >> 6, 0] + pc, son)
        assert np.all(groups[:, 1] == 0)


def test_ensemble_stimulus_groups():
    """Per-instance groups: (inst << 3) | local site combination, with the group
    table naming each instance's sites by (base, mask)."""
    from hrdayapy_amd.engine import ensemble_stim_groups
    inst = np.array([0] * 6 + [1] * 4 + [2] * 3)
    # instance 0: sites a (nodes 0,1,2), b (2,3); instance 1: none; instance 2: c (10,11)
    site_inst = [2, 0, 0]
    site_nodes = [np.array([10, 11]), np.array([0, 1, 2]), np.array([2, 3])]
    local, gs, order = ensemble_stim_groups(inst, site_inst, site_nodes, 3)
    assert list(order) == [1, 2, 0]  # sites sorted by instance, stable
    assert list(local) == [1, 1, 3, 2, 0, 0, 0, 0, 0, 0, 1, 1, 0]
    assert gs.shape == (24, 2)
    # instance 0 combination 3 = both of its sites, based at site 0
    assert tuple(gs[3]) == (0, 3)
    # instance 2's group 1 names global site base 2 (the third site after sorting)
    assert tuple(gs[8 * 2 + 1]) == (2, 1)
    with pytest.raises(ValueError):
        ensemble_stim_groups(inst, [0, 0, 0, 0], [np.array([0])] * 4, 3)
