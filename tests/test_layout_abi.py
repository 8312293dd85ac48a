"""CPU checks of the brick layout and of the C-ABI library (no GPU calls)."""
import re
from pathlib import Path

import numpy as np
import pytest

from hrdayapy_amd import _native as nat
from hrdayapy_amd.layout import BRICK_SHAPES, BrickLayout
from tests._problems import slab

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("brick", BRICK_SHAPES)
def test_layout_slots_and_neighbours(brick):
    S, _, _ = slab(13, 11, 7)
    S[4:7, 3:9, 2:5] = 0
    lay = BrickLayout(S, brick)
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
