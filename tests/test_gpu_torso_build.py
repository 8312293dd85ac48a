"""GPU torso-grid build (sharp and smoothed boundary), heart-to-torso
registration and projection, against the reference's own outputs (golden
vectors) and the oracle."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import torch
from scipy import ndimage as ndi

from hrdayapy_amd.ecg import compute_torso_grid, load_torso_grid
from hrdayapy_amd.ecg.functions import registration as REG
from hrdayapy_amd.ecg.functions import torso_build as TB
from hrdayapy_amd.ecg.functions.torso import TorsoGrid
from oracle import ecg as OE
from oracle import registration as OR
from tests._torso_host import host_grid

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(__file__)


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(os.path.join(HERE, "golden", "torso_golden.npz")))


def _gold_A(g, tag):
    n = len(g[f"{tag}_active_ids"])
    return sp.csr_matrix((g[f"{tag}_A_data"], g[f"{tag}_A_indices"], g[f"{tag}_A_indptr"]),
                         shape=(n, n))


def _smap(g):
    return {int(k): float(v) for k, v in zip(g["sigma_keys"], g["sigma_vals"])}


def test_resample_and_domain_function_match_reference(cuda, gold):
    print("[case] resample + SBM psi vs golden", flush=True)
    lab, sp_f, dx = gold["fine_labels"], gold["fine_spacing"], float(gold["dx_coarse"])
    coarse = TB.resample_labels(lab, sp_f, dx, cuda).cpu().numpy()
    assert np.array_equal(coarse, gold["coarse_labels"])
    psi, pos = TB.smoothed_domain(lab, sp_f, dx, coarse.shape, 3.0, 1e-3, cuda)
    psi = psi.cpu().numpy()
    ref = gold["psi"]
    # same cells in the system; values to FP64 round-off (tanh of the library)
    assert np.array_equal(psi > 0, ref > 0) and np.array_equal(pos.cpu().numpy() != 0, ref > 0)
    np.testing.assert_allclose(psi, ref, rtol=1e-13, atol=1e-15)
    print(f"  psi max |diff| {np.abs(psi - ref).max():.2e}", flush=True)


@pytest.mark.parametrize("spacing", [(0.75, 0.75, 1.0), (0.4, 0.5, 0.7), (1.0, 1.0, 1.0)])
def test_feature_transform_distances_match_scipy(cuda, spacing):
    print(f"[case] feature transform vs scipy EDT, spacing {spacing}", flush=True)
    rng = np.random.default_rng(11)
    m = ndi.binary_opening(rng.random((37, 29, 23)) < 0.55, iterations=1).astype(np.uint8)
    ft = TB.feature_transform(torch.from_numpy(m), 0, spacing, cuda).cpu().numpy()
    idx = np.indices(m.shape)
    d = np.sqrt(np.add.reduce(((ft - idx) * np.asarray(spacing)[:, None, None, None]) ** 2, axis=0))
    ref, rft = ndi.distance_transform_edt(m, sampling=spacing, return_indices=True)
    np.testing.assert_allclose(d, ref, rtol=1e-14, atol=0)
    # every returned site is a feature voxel; ties aside, the same site as scipy's
    assert np.all(m[ft[0], ft[1], ft[2]] == 0)
    same = np.all(ft == rft, axis=0)
    print(f"  sites equal to scipy's: {same.mean():.5f}", flush=True)


@pytest.mark.parametrize("tag", ["sharp", "sbm"])
def test_device_build_matches_reference(cuda, gold, tag):
    print(f"[case] device torso build ({tag}) vs golden stiffness", flush=True)
    g = TorsoGrid.from_segmentation(gold["fine_labels"], gold["fine_spacing"], _smap(gold),
                                    float(gold["dx_coarse"]), use_sbm=(tag == "sbm"), device=cuda)
    assert np.array_equal(g["active_ids"], gold[f"{tag}_active_ids"])
    np.testing.assert_array_equal(g["surface_flat"], gold["surface_flat"])
    if tag == "sbm":
        np.testing.assert_array_equal(g["sigma_vol"], gold["sbm_sigma"])
    A, R = g.stiffness_csr(), _gold_A(gold, tag)
    A.sort_indices()
    assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
    # sharp: the reference's arithmetic exactly; SBM: psi carries the device
    # tanh's round-off (<= 1e-13 relative), which the harmonic means pass on
    np.testing.assert_allclose(A.data, R.data, rtol=1e-14 if tag == "sharp" else 1e-12, atol=0)
    # the column-run layout equals the one read off the oracle's matrix
    psi = gold["psi"] if tag == "sbm" else None
    hg = host_grid(gold["coarse_labels"], _smap(gold), float(gold["dx_coarse"]), psi_vol=psi)
    for k in ("col_k0", "col_len", "col_off"):
        np.testing.assert_array_equal(g.layout[k], hg.layout[k], err_msg=k)
    for k in ("gx", "gy", "gz", "diag"):  # FP32 roundings of the FP64 values above
        if tag == "sharp":
            np.testing.assert_array_equal(g.layout[k], hg.layout[k], err_msg=k)
        else:
            np.testing.assert_allclose(g.layout[k], hg.layout[k], rtol=1.2e-7, err_msg=k)
    assert g.layout["pin"] == hg.layout["pin"] and g["pin_row"] == hg["pin_row"]
    np.testing.assert_array_equal(g.cell_of_row, hg.cell_of_row)


def test_full_torso_build_and_sbm(cuda):
    """The 384x384x256 torso at 1.5 mm: the device build equals the layout read
    off the oracle's matrix on a slab of columns, every row sums to zero, and the
    SBM grid from the 0.75 mm segmentation adds a band of psi < 1 cells outside
    the tissue whose conductances stay symmetric."""
    import time
    from hrdayapy_amd import geometry as G
    print("[case] full torso build, sharp and SBM", flush=True)
    lab, origin, _ = G.torso((384, 384, 256), 1.5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g = TorsoGrid.from_labels(lab, G.TORSO_SIGMA, 1.5, origin, device=cuda)
    torch.cuda.synchronize()
    print(f"  sharp build {time.perf_counter() - t0:.2f} s, {g.layout['n_cells']:,} cells",
          flush=True)
    g64 = g.g64.cpu().numpy()
    flat = g.flat_of_cell
    ni, nj, nk = g["shape"]
    cof = np.full(ni * nj * nk, -1, np.int64)
    cof[flat] = np.arange(len(flat))
    act = g["active_mask"].ravel()[flat]
    s = g64[0] + g64[1] + g64[2]
    for q, off in enumerate((nj * nk, nk, 1)):
        ok = (flat - off >= 0)
        nb = np.full(len(flat), -1)
        nb[ok] = cof[flat[ok] - off]
        has = nb >= 0
        s[has] += g64[q][nb[has]]
    row = act.copy()
    row[g.layout["pin"]] = False
    rel = np.abs(s[row] - g64[3][row]) / g64[3][row]
    assert rel.max() < 1e-14, rel.max()
    # a slab of 6 i-planes against the oracle-derived layout
    sub = lab[190:196]
    hg = host_grid(sub, G.TORSO_SIGMA, 1.5)
    gs = TorsoGrid.from_labels(sub, G.TORSO_SIGMA, 1.5, device=cuda)
    for k in ("col_k0", "col_len", "col_off", "gx", "gy", "gz", "diag"):
        np.testing.assert_array_equal(gs.layout[k], hg.layout[k], err_msg=k)
    # smoothed boundary from the 0.75 mm segmentation
    fine, _, _ = G.torso((768, 768, 512), 0.75)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gb = TorsoGrid.from_segmentation(fine, (0.75, 0.75, 0.75), G.TORSO_SIGMA, 1.5, origin,
                                     use_sbm=True, device=cuda)
    torch.cuda.synchronize()
    n_band = int((gb["active_mask"] & ~gb["tissue_mask"]).sum())
    print(f"  SBM build {time.perf_counter() - t0:.2f} s (768x768x512 segmentation), "
          f"{len(gb['active_ids']):,} rows, {n_band:,} band cells", flush=True)
    psi = gb["psi_vol"]
    # the zoom's per-axis weights sum to 1 within rounding (as scipy's)
    assert n_band > 0 and psi[gb["active_mask"]].min() >= 1e-3 and psi.max() <= 1.0 + 1e-12
    coarse = TB.resample_labels(fine, (0.75, 0.75, 0.75), 1.5, cuda).cpu().numpy()
    assert coarse.shape == (384, 384, 256)
    assert np.array_equal(gb["tissue_mask"], coarse != 0)
    assert np.all(gb["active_mask"][gb["tissue_mask"]])
    tflat = np.flatnonzero(gb["tissue_mask"].ravel())
    assert gb["pin_row"] == int(gb["active_id_map"][tflat[len(tflat) // 2]])
    gbs = gb.g64.cpu().numpy()
    assert np.all(gbs[:3] >= 0) and np.all(np.isfinite(gbs))


def test_compute_torso_grid_dropin_and_reload(cuda, gold, tmp_path):
    print("[case] compute_torso_grid / load_torso_grid (SBM)", flush=True)
    kw = dict(sigma_map=_smap(gold), dx_coarse_mm=float(gold["dx_coarse"]), verbose=False,
              fine_label_volume=gold["fine_labels"], fine_spacing=gold["fine_spacing"])
    g = compute_torso_grid(use_sbm=True, save_path=tmp_path / "tg.npz", **kw)
    for k in ("shape", "dx_mm", "origin_mm", "sigma_vol", "active_mask", "tissue_mask",
              "active_ids", "active_id_map", "surface_flat", "surface_rows", "surface_xyz",
              "pin_row", "use_sbm", "psi_vol"):
        assert k in g, k
    z = np.load(tmp_path / "tg.npz")
    assert "psi_vol" in z.files and bool(z["use_sbm"])
    # reloaded from the saved metadata (psi stored in FP32, as the reference saves it)
    g2 = load_torso_grid(tmp_path / "tg.npz")
    assert np.array_equal(g2["active_ids"], g["active_ids"]) and g2["pin_row"] == g["pin_row"]
    with pytest.raises(ValueError, match="fine segmentation"):
        compute_torso_grid(sigma_map=_smap(gold), label_volume=gold["coarse_labels"], use_sbm=True)
    gs = compute_torso_grid(sigma_map=_smap(gold), label_volume=gold["coarse_labels"],
                            dx_coarse_mm=float(gold["dx_coarse"]), verbose=False)
    A = gs["A_scipy"]
    R = _gold_A(gold, "sharp")
    assert abs(A - R).max() <= 1e-14 * abs(R).max()


def test_registration_matches_reference(cuda, gold):
    print("[case] registration (48 signed permutations) vs golden", flush=True)
    src = gold["reg_src"]
    tgt = gold["reg_tgt"].astype(np.float64)
    R, t, scale, dice = REG.register_heart_to_torso(src, tgt, 3.0, verbose=False, device=cuda)
    assert np.array_equal(R, gold["reg_R"]) and np.array_equal(t, gold["reg_t"])
    assert scale == float(gold["reg_scale"]) and dice == float(gold["reg_dice"])
    # all 48 Dice values of the search, bit for bit
    s = src * scale
    c_src = (s.max(0) + s.min(0)) / 2.0
    c_tgt = (tgt.max(0) + tgt.min(0)) / 2.0
    sub = tgt[np.random.default_rng(0).choice(len(tgt), 50_000, replace=False)]
    cnt = REG.dice_counts(s - c_src, c_tgt, sub, 3.0, cuda)
    d = [float(2 * int(c) / (int(a) + int(b))) for a, b, c in cnt]
    assert np.array_equal(np.asarray(d), gold["reg_all_dice"])
    # a finer lattice and a second cloud, against the oracle's voxel Dice
    rng = np.random.default_rng(5)
    a = rng.normal(0, 12, size=(20_000, 3))
    b = rng.normal(1, 11, size=(7_000, 3))
    cnt = REG.dice_counts(a, np.zeros(3), b, 0.7, cuda)
    for r, Rm in enumerate(OR.signed_permutations()[:8]):
        ref = OR.voxel_dice((Rm @ a.T).T, b, 0.7)
        assert float(2 * int(cnt[r, 2]) / (int(cnt[r, 0]) + int(cnt[r, 1]))) == ref


def test_projection_rows_match_reference(cuda):
    """hp_project_rows_f64 against the reference's build_projection_map output
    (mesh golden vectors) and the oracle on a rotated, scaled cloud."""
    print("[case] projection rows vs golden", flush=True)
    g = dict(np.load(os.path.join(HERE, "golden", "mesh_golden.npz")))
    from tests.test_oracle_golden import _torso
    _, _, _, idm = _torso(g)
    lab = g["torso_lab"]
    coords = g["heart_vox"][:, [1, 0, 2]].astype(np.float64) * 0.5
    a = (coords, g["proj_R"], g["proj_t"], float(g["proj_scale"]), g["proj_origin"], 1.5, idm,
         lab.shape)
    np.testing.assert_array_equal(REG.projection_rows(*a, device=cuda), g["proj_rows"])
    rng = np.random.default_rng(2)
    xyz = rng.uniform(-5, 40, size=(5000, 3))
    for Rm in OR.signed_permutations()[::7]:
        b = (xyz, Rm, np.array([3.3, -1.7, 2.9]), 1.13, np.array([0.5, -0.25, 1.0]), 1.5, idm,
             lab.shape)
        np.testing.assert_array_equal(REG.projection_rows(*b, device=cuda), OE.projection(*b))


def test_product_torso_layout_with_holes_matches_oracle_matrix(cuda):
    """An air pocket inside the tissue (inactive cells inside column runs)."""
    print("[case] device build with holes vs oracle matrix", flush=True)
    from tests.test_oracle_ecg import _torso
    lab, smap = _torso(n=(14, 12, 11))
    lab[5:7, 5:7, 3:5] = 0
    g = TorsoGrid.from_labels(lab, smap, 1.5, device=cuda)
    sig = np.zeros(lab.shape)
    for k, v in smap.items():
        sig[lab == k] = v
    A, ids, idm = OE.assemble_stiffness(sig, 1.5e-3, lab != 0)
    assert np.array_equal(g["active_ids"], ids)
    assert g["pin_row"] == OE.pin_row_of(lab != 0, idm)
    assert np.array_equal(g["surface_flat"], OE.surface_flat(lab, lab != 0))
    Ag = g.stiffness_csr()
    assert abs(Ag - A).max() <= 1e-14 * abs(A).max()
    hg = host_grid(lab, smap, 1.5)
    for k in ("col_k0", "col_len", "col_off", "gx", "gy", "gz", "diag"):
        np.testing.assert_array_equal(g.layout[k], hg.layout[k], err_msg=k)
