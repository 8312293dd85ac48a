"""GPU checks of the surrogate's MFMA kernels: GEMM against numpy, gradients
against CPU autograd, fused inference against a CPU forward, and the learned
model against the same-cost alternative without it (one 0.2 ms Rush-Larsen
step in place of ten 0.02 ms substeps)."""
import numpy as np
import pytest
import torch

from hrdayapy_amd import _native as nat
from hrdayapy_amd import surrogate as SG

pytestmark = pytest.mark.gpu


def _bf(x):
    return torch.as_tensor(x).to(torch.bfloat16)


@pytest.mark.parametrize("M,N,K,trans_a,trans_b,ksplit", [
    (64, 64, 32, False, False, 1), (100, 37, 70, False, True, 1), (19, 128, 3000, True, False, 8),
    (1000, 128, 22, False, True, 1)])
def test_gemm_matches_numpy(cuda, M, N, K, trans_a, trans_b, ksplit):
    print(f"[case] gemm {M}x{N}x{K} ta={trans_a} tb={trans_b} ks={ksplit}", flush=True)
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((K, M) if trans_a else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if trans_b else (K, N)).astype(np.float32)
    Ab, Bb = _bf(A), _bf(B)
    Ar, Br = Ab.float().numpy(), Bb.float().numpy()
    ref = (Ar.T if trans_a else Ar) @ (Br.T if trans_b else Br)
    Ad, Bd = Ab.to(cuda), Bb.to(cuda)
    C = torch.zeros((M, N), dtype=torch.float32, device=cuda)
    sam, sak = (1, M) if trans_a else (K, 1)
    sbk, sbn = (1, K) if trans_b else (N, 1)
    nat.call("hp_gemm_bf16", M, N, K, Ad.data_ptr(), sam, sak, Bd.data_ptr(), sbk, sbn,
             C.data_ptr(), N, 1.0, 0.0, None, ksplit, nat.stream_handle())
    got = C.cpu().numpy()
    assert np.max(np.abs(got - ref)) < 1e-3 * max(1.0, np.abs(ref).max())


def _cpu_forward(net, Xn, bf16_acts=True):
    """FP32 CPU forward with the same bf16 rounding points as the kernels."""
    W = [w.to(torch.bfloat16).float().cpu() for w in net.W]
    b = [x.cpu() for x in net.b]
    a = Xn.float().cpu()[:, :SG.N_IN]
    zs = []
    for l in range(3):
        z = a @ W[l].t() + b[l]
        zs.append(z)
        a = torch.nn.functional.silu(z)
        if bf16_acts:
            a = a.to(torch.bfloat16).float()
    return a @ W[3].t() + b[3], zs


def test_gradients_match_cpu_autograd(cuda):
    print("[case] surrogate gradients", flush=True)
    torch.manual_seed(0)
    net = SG.SurrogateNet(cuda, seed=3)
    for b in net.b:
        b.normal_(0, 0.1)
    net.refresh()
    B = 512
    tr = SG.Trainer(net, batch=B)
    Xn = torch.zeros((B, 32), dtype=torch.bfloat16, device=cuda)
    Xn[:, :SG.N_IN] = torch.randn(B, SG.N_IN, device=cuda).to(torch.bfloat16)
    Tn = torch.randn(B, SG.N_OUT, device=cuda)
    Y = tr.forward(Xn, B).clone()
    tr.backward(Xn, Tn, B)
    # CPU autograd on the same bf16 weights (parameters as leaves)
    W = [w.to(torch.bfloat16).float().cpu().requires_grad_() for w in net.W]
    bb = [x.cpu().clone().requires_grad_() for x in net.b]
    a = Xn.float().cpu()[:, :SG.N_IN]
    for l in range(3):
        a = torch.nn.functional.silu(a @ W[l].t() + bb[l])
    Yc = a @ W[3].t() + bb[3]
    loss = ((Yc - Tn.cpu()) ** 2).mean()
    loss.backward()
    assert torch.allclose(Y.cpu(), Yc.detach(), rtol=3e-2, atol=3e-2)
    for l in range(4):
        g_gpu = tr.gW[l].cpu()
        g_cpu = W[l].grad
        rel = float((g_gpu - g_cpu).norm() / g_cpu.norm())
        print(f"  layer {l}: relative gradient error {rel:.2e}")
        assert rel < 5e-2
        relb = float((tr.gb[l].cpu() - bb[l].grad).norm() / bb[l].grad.norm())
        assert relb < 5e-2


@pytest.mark.parametrize("variant,listed,correct", [
    (0, False, 0), (1, False, 0), (1, True, 0), (1, "state", 0), (2, "state", 0), (0, False, 1),
    (1, "state", 1), (2, "state", 1), (2, "state", 2)])
def test_inference_kernel_matches_cpu_forward(cuda, variant, listed, correct):
    """correct = 1: the network's output is added to the state and rate a big
    step left (fallback voxels are the ones with rate NaN, voxels at rest keep
    the big step; V_prev untouched); correct = 2: also a trust bound at the
    median |V correction|, above which a voxel keeps the big step."""
    print(f"[case] fused surrogate inference, kernel variant {variant}, slot list {listed}, "
          f"correction mode {correct}", flush=True)
    from hrdayapy_amd.layout import BrickLayout
    S = np.zeros((12, 10, 9), bool)
    S[1:11, 1:9, 1:8] = True
    lay = BrickLayout(S, (4, 4, 4))
    n = lay.n_nodes
    rng = np.random.default_rng(7)
    net = SG.SurrogateNet(cuda, seed=5)
    for b in net.b:
        b.normal_(0, 0.05)
    # normalisation from a random state cloud around rest/plateau
    V = torch.from_numpy(rng.uniform(-88, 20, n)).float()
    st = torch.tensor([[SG.INIT[k] for k in SG.STATE_NAMES]]).repeat(n, 1).float()
    st *= torch.from_numpy(rng.uniform(0.8, 1.2, (n, 18))).float()
    ct = torch.from_numpy(rng.integers(0, 3, n))
    X = SG.SurrogateNet.features(V, st, ct)
    Xt = SG.SurrogateNet.transform(X)
    D = torch.randn(n, SG.N_OUT) * 0.01
    net.fit_normalisation(Xt, D)
    # engine-like buffers
    Vd = torch.from_numpy(lay.to_slots(V.numpy())).to(cuda)
    Vprev = Vd.clone()
    Vprev[torch.from_numpy(lay.slot_of_node[:10].astype(np.int64)).to(cuda)] -= 5.0  # fallback
    stc = st.clone()
    stc[:, 0] -= SG.KI_BASE
    stc[:, 1] -= SG.NAI_BASE
    # compact state: entry state_of_node[i] holds node i
    groups, son = lay.state_groups()
    son_t = torch.from_numpy(son.astype(np.int64))
    state = torch.zeros((18, n + 64), dtype=torch.float32)
    state[:, son_t] = stc.t()
    state = state.to(cuda).contiguous()
    state0 = state.clone()
    lab = torch.from_numpy(lay.labels(ct.numpy().astype(np.uint8))).to(cuda)
    nbr = torch.from_numpy(lay.nbr).to(cuda)
    grp = torch.from_numpy(groups).to(cuda)
    br = nat.HpBricks(lay.n_bricks, 4, 4, 4, 0, nbr.data_ptr(), lab.data_ptr(), None, None,
                      grp.data_ptr())
    sl = torch.from_numpy(lay.slot_of_node.astype(np.int64))
    Vn = V.numpy()
    fb = np.zeros(n, bool)
    fb[:10] = True
    fb |= (Vn > -70) & (Vn < -40)
    rate = torch.zeros(lay.n_slots, device=cuda)
    r0 = np.zeros(n, np.float32)
    if correct:  # the big step's rate: a number, NaN on its fallback voxels
        r0 = rng.uniform(-2.0, 2.0, n).astype(np.float32)
        r0[fb] = np.nan
        rate[sl.to(cuda)] = torch.from_numpy(r0).to(cuda)
    Vprev0 = Vprev.clone()
    sp = nat.HpSurrogateParams(0.2, 1.0, -70.0, -40.0, SG.KI_BASE, SG.NAI_BASE, 0, variant)
    sp.correct = min(correct, 1)
    # CPU forward (bf16 inputs, as the kernel rounds them)
    Xn = ((Xt - torch.from_numpy(net.in_mean)) / torch.from_numpy(net.in_std)).to(torch.bfloat16)
    Yc, _ = _cpu_forward(net, Xn)
    Dc = Yc * torch.from_numpy(net.out_std) + torch.from_numpy(net.out_mean)
    rest = np.zeros(n, bool)
    near = np.zeros(n, bool)  # V correction within 5 % of the trust bound: not checked
    if correct:  # voxels at rest keep the big step's state and rate
        sp.rest_v, sp.rest_rate, sp.corr_max = -70.0, 1.0, 1e30
        rest = (Vn < -70.0) & (np.abs(r0) < 1.0) & ~fb
        assert rest.sum() > 20
    if correct == 2:  # so do voxels whose V correction exceeds the bound
        a = np.abs(Dc[:, 0].numpy())
        sp.corr_max = float(np.median(a[~fb & ~rest]))
        near = np.abs(a - sp.corr_max) < 0.05 * sp.corr_max
        rest |= (a > sp.corr_max) & ~fb & ~near
    if listed:  # the tissue slots in a shuffled order, padded with -1 to a multiple of 32
        perm = rng.permutation(lay.n_nodes)
        tis = lay.slot_of_node[perm].astype(np.int32)
        lst = np.full((len(tis) + 31) // 32 * 32, -1, np.int32)
        lst[:len(tis)] = tis
        lst_d = torch.from_numpy(lst).to(cuda)
        sp.slot_list, sp.n_list = lst_d.data_ptr(), len(lst)
        if listed == "state":  # with the state index of each listed slot
            stl = np.zeros_like(lst)
            stl[:len(tis)] = son[perm]
            stl_d = torch.from_numpy(stl).to(cuda)
            sp.state_list = stl_d.data_ptr()
    import ctypes
    nat.call("hp_surrogate_step", ctypes.byref(br), Vd.data_ptr(), Vprev.data_ptr(),
             state.data_ptr(), n + 64, rate.data_ptr(), ctypes.byref(net.hp_w),
             ctypes.byref(net.hp_norm), ctypes.byref(sp), 4, nat.stream_handle())
    r = rate.cpu()[sl]
    assert np.array_equal(np.isnan(r.numpy()), fb)
    if correct:
        assert torch.equal(Vprev, Vprev0)  # the big step's V_prev is kept
    else:
        assert torch.equal(Vprev, Vd)
    ok = ~fb & ~rest & ~near
    np.testing.assert_allclose(r.numpy()[ok], r0[ok] + (Dc[:, 0] / 0.2).numpy()[ok], rtol=5e-2,
                               atol=2e-3)
    assert np.array_equal(r.numpy()[rest], r0[rest])
    new = state.cpu()[:, son_t].t()
    old = state0.cpu()[:, son_t].t()
    # untouched on fallback voxels and, in correction mode, on voxels at rest
    # and beyond the trust bound
    keep = torch.from_numpy(fb | rest)
    assert torch.equal(new[keep], old[keep])
    # updated as f^-1(f(x) + d) elsewhere
    lg = torch.from_numpy(SG.IN_LOG[1:19].astype(bool))
    oa = old.clone()
    oa[:, 0] += SG.KI_BASE
    oa[:, 1] += SG.NAI_BASE
    exp_new = oa + Dc[:, 1:]
    exp_new[:, lg] = torch.exp(torch.log(oa[:, lg]) + Dc[:, 1:][:, lg])
    na = new.clone()
    na[:, 0] += SG.KI_BASE
    na[:, 1] += SG.NAI_BASE
    okt = torch.from_numpy(ok)
    rel = ((na[okt] - exp_new[okt]).abs() / (exp_new[okt].abs() + 1e-9)).max()
    assert float(rel) < 2e-2


@pytest.fixture(scope="module")
def trained(cuda):
    """Network trained (HIP kernels) on the plain integrator's trajectories of a
    transmural slab, as the bench does at full size."""
    eng = SG.training_slab(cuda, nx=40, ny=40, nz=12)
    X, D = SG.build_training_set(eng, 2000, 1500, seed=1)
    net, tr, st = SG.train_surrogate(X, D, cuda, epochs=10, batch=2048, lr_decay=0.85, log=print)
    print(f"  {st['n_train']} samples, {st['train_s']:.1f} s, held-out |dV err| "
          f"mean {st['heldout_dv_mae']:.2e} max {st['heldout_dv_max']:.2e} mV")
    return net, st


def test_training_converges(cuda, trained):
    print("[case] surrogate training", flush=True)
    net, st = trained
    assert st["loss"][-1] < 0.25 * st["loss"][0]
    assert st["heldout_dv_mae"] < 0.05  # mV over a 0.2 ms macro-step
    # the corrected macro-step is closer to the ten substeps than the big step
    print(f"  held-out |dV err|: big step {st['heldout_bigstep_dv_mae']:.2e}, "
          f"corrected {st['heldout_dv_mae']:.2e} mV")
    assert st["heldout_dv_mae"] < 0.5 * st["heldout_bigstep_dv_mae"]


def _slab_run(cuda, form, net=None, T=120.0, compact=True, graph=False):
    from hrdayapy_amd.anisotropic_simulation.functions.purkinje_myocardium_anisotropic_solver import (
        build_coupled_system)
    from tests._problems import corner_region, slab, stub_tree
    S, phi, f = slab(40, 32, 10)
    nodes, el, act = stub_tree(S, 0.25)
    kw = dict(voxel_size=0.25, dx_p=0.25, sigma_P=0.0, sigma_l=0.24, sigma_t=0.08, c_pmj=0.0,
              T=T, ectopic_region=corner_region(S), phi_endo_max=1 / 3, phi_epi_min=2 / 3,
              stim_protocol=[dict(name="c", target="ectopic", amp=0.52, dur=2.0, onset=0.0)])
    eng, meta = build_coupled_system(nodes, el, act, S, phi, f, dt=0.02, device=cuda, **kw)
    n_macro = int(T / 0.2)
    trace = []
    if form == "plain":
        for k in range(n_macro):
            eng.run(SG.N_SUB, graph_chunk=SG.N_SUB)
            if (k + 1) % 10 == 0:
                trace.append(eng.V_myo_nodes().double().cpu().numpy())
        fb = 1.0
    else:
        run = SG.SurrogateRunner(eng, net, mode=form, compact=compact)
        fbs = []
        done = 0
        while done < n_macro:
            k = 2 if graph else 1  # the graph replays two macro-steps at once
            run.run(k, graph=graph)
            fbs.append(run.fallback_fraction())
            done += k
            if done % 10 == 0:
                trace.append(eng.V_myo_nodes().double().cpu().numpy())
        fb = float(np.mean(fbs))
    return eng.at_nodes().double().cpu().numpy(), np.array(trace), fb


def test_surrogate_forms_vs_plain(cuda, trained):
    """Activation times and Vm traces of the multirate forms against the plain
    FP32 integrator on a transmural slab (120 ms: upstroke, plateau, early
    repolarisation)."""
    print("[case] surrogate forms in tissue", flush=True)
    net, _ = trained
    at0, tr0, _ = _slab_run(cuda, "plain")
    res = {}
    for form in ("mlp", "bigstep"):
        at, tr, fb = _slab_run(cuda, form, net)
        both = np.isfinite(at) & np.isfinite(at0)
        err = float(np.max(np.abs(at[both] - at0[both])))
        rms = float(np.sqrt(np.mean((tr - tr0) ** 2)))
        res[form] = (err, rms, fb)
        print(f"  {form}: AT max err {err:.3f} ms, Vm RMS {rms:.3f} mV, "
              f"mean fallback fraction {fb:.3f}, activated {both.sum()}/{len(at)}")
        assert both.sum() == np.isfinite(at0).sum()
        assert err <= 1.0
    assert res["mlp"][2] < 0.5
    # what the network gains: its correction brings the big step's Vm closer
    # to the plain integrator's
    assert res["mlp"][1] < res["bigstep"][1]


@pytest.mark.parametrize("form", ["bigstep", "mlp"])
def test_compacted_fallback_matches_divergent(cuda, trained, form):
    """The compacted fallback path (list + hp_fallback_react + diffusion-only
    kernel, replayed as a HIP graph) against the divergent path (fused kernel
    with do_reaction = 2, launched eagerly): the fallback voxels run the same
    cell step in both, so Vm and activation times agree to round-off."""
    print(f"[case] compacted fallback {form}", flush=True)
    net, _ = trained
    at0, tr0, fb0 = _slab_run(cuda, form, net, T=40.0, compact=False)
    at1, tr1, fb1 = _slab_run(cuda, form, net, T=40.0, compact=True, graph=True)
    dv = float(np.max(np.abs(tr1 - tr0)))
    both = np.isfinite(at0) & np.isfinite(at1)
    dat = float(np.max(np.abs(at1[both] - at0[both])))
    print(f"  max |dV| {dv:.2e} mV, max |dAT| {dat:.2e} ms, fallback {fb0:.3f} / {fb1:.3f}")
    assert np.array_equal(np.isfinite(at0), np.isfinite(at1))
    assert dv < 1e-2 and dat < 1e-3


def _config5_run(cuda, form, ecg, net=None, T=300.0):
    """The configuration-1 slab (64x64x16, all endocardial) for T ms in the form
    given, 12 leads every 1 ms through the lead field; returns (AT, leads,
    mean fallback fraction)."""
    from hrdayapy_amd import workloads as W
    from hrdayapy_amd.anisotropic_simulation.functions.purkinje_myocardium_anisotropic_solver import (
        build_coupled_system)
    from tests.test_gpu_configs import _config1
    args, kw = _config1(T)
    dtype = torch.float64 if form == "ref64" else torch.float32
    solver = "bicgstab" if form == "ref64" else "jacobi"
    eng, meta = build_coupled_system(*args, dt=0.02, dtype=dtype, device=cuda, scheme="cn",
                                     cn_solver=solver, **kw)
    if ecg is None:
        ecg = W.EcgSetup(eng, meta, n_bspm=0, log=lambda m: None)
    rec = W.EcgRecorder(ecg.lf, eng)
    run = None if form in ("ref64", "plain") else SG.SurrogateRunner(eng, net, mode=form,
                                                                      graph_macros=5)
    fbs = []
    for _ in range(int(T)):
        if run is None:
            eng.run(50, graph_chunk=50)
        else:
            run.run(5)
            fbs.append(run.fallback_fraction())
        rec.record()
    phi = rec.potentials()
    leads = W.twelve_leads(ecg.el.names, phi)
    return eng.at_nodes().double().cpu().numpy(), leads, (float(np.mean(fbs)) if fbs else 1.0), ecg


@pytest.mark.timeout(900)
def test_config5_surrogate_on_slab_with_torso_ecg(cuda):
    """Configuration 5: the network trained by the build's own kernels on ~50 M
    transitions of the slab runs (the reference's protocol and a delayed-onset
    run), evaluated on the configuration-1 slab with the torso ECG (12 leads
    every 1 ms for 300 ms) against the reference's scheme in FP64 (CN +
    BiCGSTAB), beside the network-free big step with the same fallback rule
    and the plain FP32 form.

    What the network gains over the method without it: on held-out
    transitions its corrected macro-step is ~27 times closer to the ten exact
    substeps than the big step alone (measured 2.8e-5 against 7.7e-4 mV mean
    |dV error|), and in the slab its leads are closer to the reference's than
    the big step's.  Both multirate forms meet the activation-time gate
    (0.04 ms against 1 ms); on this small heart both miss the 0.999 lead gate
    in V5/V6 (r = 0.9983, the T wave), where the plain form meets it (0.99902,
    0.03 ms): the gap
    is the multirate splitting (the reaction held over 0.2 ms while diffusion
    moves V), which a 27 times smaller reaction error leaves unchanged."""
    from hrdayapy_amd import workloads as W
    from hrdayapy_amd.ecg.functions.forward import LEAD_NAMES
    print("[case] configuration 5: surrogate on the slab with torso ECG", flush=True)
    X, D = SG.training_runs(cuda, samples=50_000_000)
    net, _, st = SG.train_surrogate(X, D, cuda, epochs=14)
    del X, D
    torch.cuda.empty_cache()
    print(f"  trained on {st['n_train']:,} transitions in {st['train_s']:.1f} s, held-out "
          f"|dV err| mean {st['heldout_dv_mae']:.2e} max {st['heldout_dv_max']:.2e} mV (big step "
          f"alone: mean {st['heldout_bigstep_dv_mae']:.2e} max {st['heldout_bigstep_dv_max']:.2e})",
          flush=True)
    at0, l0, _, ecg = _config5_run(cuda, "ref64", None)
    res = {}
    for form in ("plain", "mlp", "bigstep"):
        at, leads, fb, _ = _config5_run(cuda, form, ecg, net)
        both = np.isfinite(at) & np.isfinite(at0)
        err = float(np.max(np.abs(at[both] - at0[both])))
        r, rel, per = W.lead_agreement(leads, l0, per_lead=True)
        res[form] = (err, r, rel, fb, int(both.sum()))
        print(f"  {form}: AT max err {err:.3f} ms over {both.sum():,} voxels, 12 leads min r "
              f"{r:.6f}, max rel L2 {rel:.3e}, mean fallback fraction {fb:.3f}", flush=True)
        print("    per lead (r, rel L2): " + ", ".join(
            f"{n} {a}/{b}" for n, (a, b) in zip(LEAD_NAMES, per)), flush=True)
    for form, (err, r, rel, fb, n) in res.items():
        assert n == len(at0) and err <= 1.0  # every voxel activates, AT gate met
        assert form == "plain" or fb < 0.5
    assert res["plain"][1] >= 0.999  # the plain form meets the lead gate
    # the gain: per macro-step on held-out transitions, and in the leads
    assert st["heldout_dv_mae"] < 0.1 * st["heldout_bigstep_dv_mae"]
    assert res["mlp"][2] < res["bigstep"][2] and res["mlp"][1] >= res["bigstep"][1]
    assert res["mlp"][1] >= 0.998
