"""Batched ensemble (one brick list, per-instance conductivity, g_Ks/g_to and
pacing sites) against the FP64 oracle per instance, and against the same
instances run one at a time."""
import numpy as np
import pytest
import torch

from hrdayapy_amd import ensemble as E
from hrdayapy_amd import geometry as G
from oracle import monodomain as OM
from tests._problems import slab, stub_tree

pytestmark = pytest.mark.gpu


def _site(pick, onset, radius_mm=1.0):
    return dict(pick=pick, onset=onset, dur=2.0, amp=0.52, radius_mm=radius_mm)


def _instances():
    """Three small instances of different shape, voxel size and parameters."""
    S1, p1, f1 = slab(20, 16, 8)
    S2, p2, f2 = slab(14, 22, 6, fibre="x", seed=3)
    geo = G.biv_shell(40, 0.4, seed=2)
    geoms = [(S1, p1, f1, 0.25), (S2, p2, f2, 0.25),
             (geo["S"], geo["phi"], geo["fibre_dir"], 0.4)]
    specs = [
        E.InstanceSpec("slab", 0, 0, 0.24 * 1.25, 0.08 * 0.75, 1.45, 0.55,
                       [_site(0.1, 0.0), _site(0.8, 6.0)]),
        E.InstanceSpec("slab", 0, 0, 0.24 * 0.8, 0.08 * 1.2, 0.6, 1.4, [_site(0.5, 2.0)]),
        E.InstanceSpec("biv", 0, 0, 0.24, 0.08, 1.0, 1.0,
                       [_site(0.2, 0.0), _site(0.6, 1.0), _site(0.9, 3.0)]),
    ]
    return specs, geoms


def _oracle_instance(spec, geom, site_ids, T, scheme="explicit"):
    """The instance alone through the oracle (FP64)."""
    S, phi, f, h = geom
    nodes, el, act = stub_tree(S, h)
    vox = np.argwhere(np.asarray(S).astype(bool))
    regions, proto = {}, []
    for k, (s, ids) in enumerate(zip(spec.sites, site_ids)):
        m = np.zeros(S.shape, bool)
        m[tuple(vox[ids].T)] = True
        regions[f"s{k}"] = m
        proto.append(dict(name=f"s{k}", target="region", region=f"s{k}", amp=s["amp"],
                          dur=s["dur"], onset=s["onset"]))
    prob = OM.Problem(nodes, el, act, S, phi, f, voxel_size=h, dx_p=h, sigma_P=0.0,
                      sigma_l=spec.sigma_l, sigma_t=spec.sigma_t, c_pmj=0.0, T=T,
                      stim_protocol=proto, stim_regions=regions, phi_endo_max=1 / 3,
                      phi_epi_min=2 / 3)
    gks = np.concatenate([np.ones(prob.Np), np.full(prob.N_myo, spec.gks_scale)])
    gto = np.concatenate([np.ones(prob.Np), np.full(prob.N_myo, spec.gto_scale)])
    r = OM.run(prob, T, 0.02, scheme=scheme, g_ks_scale=gks, g_to_scale=gto)
    return r["V"][prob.Np:], r["at"][prob.Np:]


def _site_ids(ens, i):
    lo = ens.node_offset[i]
    return [s["node_ids"] - lo for s in ens.sites if s["inst"] == i]


def test_ensemble_f64_matches_oracle(cuda):
    print("[case] ensemble f64 vs oracle", flush=True)
    specs, geoms = _instances()
    T = 24.0
    ens = E.Ensemble(specs, geoms=geoms, dtype=torch.float64, device=cuda, brick=(4, 4, 4))
    ens.run(int(round(T / 0.02)), graph_chunk=50)
    Vs, ats = ens.V_instances(), ens.at_instances()
    for i, (spec, geom) in enumerate(zip(specs, geoms)):
        Vo, ato = _oracle_instance(spec, geom, _site_ids(ens, i), T)
        both = np.isfinite(ato) & np.isfinite(ats[i])
        err_v = float(np.max(np.abs(Vs[i] - Vo)))
        err_at = float(np.max(np.abs(ats[i][both] - ato[both]))) if both.any() else 0.0
        print(f"  instance {i} ({spec.kind}, {len(Vo)} voxels): max |dV| {err_v:.2e} mV, "
              f"max |dAT| {err_at:.2e} ms, activated {both.sum()}")
        assert np.array_equal(np.isfinite(ato), np.isfinite(ats[i]))
        assert both.sum() > 0
        assert err_v < 1e-7 and err_at < 1e-7


def test_ensemble_cn_matches_oracle_per_instance(cuda):
    """scheme "cn": the batch, stopped on its whole residual at the batch
    tolerance, against each instance alone through the oracle's CN step
    (BiCGSTAB to 1e-5 of that instance's own system), FP32 and FP64."""
    specs, geoms = _instances()
    T = 24.0
    ref = None
    for dtype in (torch.float64, torch.float32):
        print(f"[case] ensemble cn {dtype} vs oracle", flush=True)
        ens = E.Ensemble(specs, geoms=geoms, dtype=dtype, device=cuda, brick=(4, 4, 4),
                         scheme="cn")
        ens.run(int(round(T / 0.02)), graph_chunk=50)
        st = ens.eng.cn_stats()
        print(f"  batch tol {ens.batch_tol:.2e}: {st}")
        assert st["not_converged"] == 0
        Vs, ats = ens.V_instances(), ens.at_instances()
        if ref is None:
            ref = [_oracle_instance(spec, geom, _site_ids(ens, i), T, scheme="cn")
                   for i, (spec, geom) in enumerate(zip(specs, geoms))]
        for i, (Vo, ato) in enumerate(ref):
            both = np.isfinite(ato) & np.isfinite(ats[i])
            rms = float(np.sqrt(np.mean((Vs[i] - Vo) ** 2)))
            err_at = float(np.max(np.abs(ats[i][both] - ato[both])))
            print(f"  instance {i} ({specs[i].kind}, {len(Vo)} voxels): Vm RMS {rms:.2e} mV, "
                  f"max |dAT| {err_at:.2e} ms, activated {both.sum()}")
            assert np.array_equal(np.isfinite(ato), np.isfinite(ats[i]))
            assert both.sum() > 0
            assert rms < 0.5 and err_at < 0.5


def test_ensemble_f32_equals_instances_alone(cuda):
    """Instances share one brick list but never exchange current: each one's
    FP32 result equals its run alone, bit for bit."""
    print("[case] ensemble f32 vs instances alone", flush=True)
    specs, geoms = _instances()
    n = 600
    ens = E.Ensemble(specs, geoms=geoms, dtype=torch.float32, device=cuda, brick=(4, 4, 4))
    ens.run(n, graph_chunk=50)
    Vs, ats = ens.V_instances(), ens.at_instances()
    for i in range(len(specs)):
        one = E.Ensemble([specs[i]], geoms=[geoms[i]], dtype=torch.float32, device=cuda,
                         brick=(4, 4, 4))
        one.run(n, graph_chunk=50)
        V1, at1 = one.V_instances()[0], one.at_instances()[0]
        assert np.array_equal(Vs[i], V1), i
        assert np.array_equal(np.isnan(ats[i]), np.isnan(at1))
        ok = ~np.isnan(at1)
        assert np.array_equal(ats[i][ok], at1[ok])


def test_ensemble_parameters_matter(cuda):
    """Sanity of the parameter block: the same geometry with g_Ks x0.5 vs x1.5
    (everything else equal) must give different repolarisation."""
    print("[case] ensemble parameter effect", flush=True)
    S, p, f = slab(12, 12, 6)
    g = (S, p, f, 0.25)
    a = E.InstanceSpec("slab", 0, 0, 0.24, 0.08, 0.5, 1.0, [_site(0.0, 0.0, 2.0)])
    b = E.InstanceSpec("slab", 0, 0, 0.24, 0.08, 1.5, 1.0, [_site(0.0, 0.0, 2.0)])
    ens = E.Ensemble([a, b], geoms=[g, g], dtype=torch.float32, device=cuda, brick=(4, 4, 4))
    ens.run(5000, graph_chunk=50)  # 100 ms: plateau, where I_Ks acts
    Va, Vb = ens.V_instances()
    print(f"  mean V at 100 ms: g_Ks x0.5 {Va.mean():.2f} mV, x1.5 {Vb.mean():.2f} mV")
    assert Va.mean() > Vb.mean() + 0.5
