"""Learned TTP06 surrogate: data generation with the plain integrator, training
with the project's own HIP kernels (BF16 MFMA GEMMs, SiLU, MSE, Adam), and
the surrogate form of the monodomain time loop.

The network maps the cell state at the start of a 0.2 ms macro-step
(V, the 18 TTP06 variables, phenotype one-hot) to the increments that ten
0.02 ms reaction substeps (Rush-Larsen gates, forward-Euler concentrations,
V += -dt*I_ion) produce.  In tissue, voxels at rest or on the plateau take the
network's increment once per macro-step (V through a constant rate while
diffusion still runs every substep); voxels whose |dV/dt| exceeds 1 mV/ms or
whose V lies in (-70, -40) mV fall back to the plain integrator.

Two roles of the network (``SurrogateNet.correct``):

- direct: it predicts the whole increment of the ten substeps;
- correction (the default of ``train_surrogate``): one 0.2 ms step of the
  exact cell model (the big step) is taken first, and the network, fed V and
  the advanced state, predicts what the ten substeps add to it.  The big
  step's own error is ~200 times smaller than the increment, so the same
  network trained to the same relative accuracy leaves an error about that
  much smaller; what the network gains is measured against the big step
  alone (the same rule without the network).
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import torch

from . import _native as nat
from .ttp06 import INIT, STATE_NAMES

N_IN, N_HID, N_OUT = 22, 128, 19
# features carried as logarithms: Cai, m, d, r, Cass, Xs, Xr1 (span orders of magnitude)
LOG_FEATURES = ("Cai", "m", "d", "r", "Cass", "Xs", "Xr1")
FEATURE_NAMES = ["V"] + STATE_NAMES + ["is_endo", "is_mid", "is_epi"]
IN_LOG = np.array([1 if n in LOG_FEATURES else 0 for n in FEATURE_NAMES], dtype=np.int32)
KI_BASE, NAI_BASE = INIT["Ki"], INIT["Nai"]
T_MACRO, N_SUB = 0.2, 10
# correction mode: voxels at rest (V < REST_V mV and |dV/dt| < REST_RATE mV/ms
# by the big step's own rate) keep the big step alone: there its error (~3e-6
# mV per macro-step on paced single cells) is below the network's
REST_V, REST_RATE = -80.0, 0.05
# correction mode: a V correction above CORR_MAX mV per macro-step is not
# trusted (the voxel keeps the big step alone); the big step's own error is
# below ~0.1 mV on paced single cells
CORR_MAX = 0.5
# output loss weights of train_surrogate (see SurrogateNet.fit_normalisation)
LOSS_WEIGHTS = {"V": 4.0, "h": 3.0, "j": 3.0, "m": 2.0}


def _bf16(t):
    return t.to(torch.bfloat16).contiguous()


class SurrogateNet:
    """FP32 master weights + BF16 copies on the GPU, and the normalisation."""

    def __init__(self, device="cuda", seed=0):
        g = torch.Generator().manual_seed(seed)
        dev = torch.device(device)
        self.dev = dev

        def init(o, i):
            return (torch.randn(o, i, generator=g) * math.sqrt(2.0 / i)).to(dev)

        self.W = [init(N_HID, N_IN), init(N_HID, N_HID), init(N_HID, N_HID),
                  init(N_OUT, N_HID) * 0.1]
        self.b = [torch.zeros(N_HID, device=dev), torch.zeros(N_HID, device=dev),
                  torch.zeros(N_HID, device=dev), torch.zeros(N_OUT, device=dev)]
        self.in_mean = np.zeros(N_IN, np.float32)
        self.in_std = np.ones(N_IN, np.float32)
        self.out_mean = np.zeros(N_OUT, np.float32)
        self.out_std = np.ones(N_OUT, np.float32)
        self.correct = False  # True: the output corrects the big step (module docstring)
        self.refresh()

    # -- parameters as C structs
    def refresh(self):
        self.Wb = [_bf16(w) for w in self.W]
        self.refresh_inference()
        self._refresh_norm()

    def refresh_inference(self):
        """The inference kernel's scaled weights (hp_mlp_weights): w1, b1..b3
        times -log2(e), w4 times -ln 2, from the FP32 master weights.  Call
        after the master weights change (training)."""
        c = 1.0 / math.log(2.0)
        self.Wi = [_bf16(self.W[0] * -c), _bf16(self.W[1]), _bf16(self.W[2]),
                   _bf16(self.W[3] * -math.log(2.0))]
        self.bi = [(b * -c).contiguous() for b in self.b[:3]] + [self.b[3].clone()]
        self.hp_w = nat.HpMlpWeights(*[w.data_ptr() for w in self.Wi],
                                     *[b.data_ptr() for b in self.bi])

    def _refresh_norm(self):
        n = nat.HpMlpNorm()
        for k in range(N_IN):
            n.in_mean[k] = float(self.in_mean[k])
            n.in_rstd[k] = float(1.0 / self.in_std[k])
            n.in_log[k] = int(IN_LOG[k])
        for k in range(N_OUT):
            n.out_mean[k] = float(self.out_mean[k])
            n.out_std[k] = float(self.out_std[k])
        self.hp_norm = n

    def state_dict(self):
        return {"W": [w.cpu() for w in self.W], "b": [b.cpu() for b in self.b],
                "in_mean": self.in_mean, "in_std": self.in_std, "out_mean": self.out_mean,
                "out_std": self.out_std, "correct": bool(self.correct)}

    def load_state_dict(self, d):
        self.W = [torch.as_tensor(w).to(self.dev) for w in d["W"]]
        self.b = [torch.as_tensor(b).to(self.dev) for b in d["b"]]
        for k in ("in_mean", "in_std", "out_mean", "out_std"):
            setattr(self, k, np.asarray(d[k], np.float32))
        self.correct = bool(d.get("correct", False))
        self.refresh()

    # -- feature transforms (host side of data preparation)
    @staticmethod
    def features(V, S, ct):
        """(n,22) raw features: V, 18 absolute state variables, one-hot phenotype."""
        n = V.shape[0]
        X = torch.empty((n, N_IN), dtype=torch.float32, device=V.device)
        X[:, 0] = V
        X[:, 1:19] = S
        X[:, 19:] = torch.nn.functional.one_hot(ct.long(), 3).float()
        return X

    @staticmethod
    def transform(X):
        X = X.clone()
        lg = torch.from_numpy(IN_LOG.astype(bool)).to(X.device)
        X[:, lg] = torch.log(X[:, lg].clamp_min(1e-30))
        return X

    def fit_normalisation(self, Xt, Dt, weights=None):
        """Per-feature mean/std of the inputs and outputs.  ``weights`` maps output
        names (FEATURE_NAMES[:19]: "V" and the state variables) to loss weights w:
        the output's scale is divided by w, so its error counts w^2 times more in
        the MSE on normalised outputs (V, and the Na gates h and j that set the
        conduction velocity, are what the tissue and the ECG see)."""
        self.in_mean = Xt.mean(0).cpu().numpy().astype(np.float32)
        sd = Xt.std(0).cpu().numpy().astype(np.float32)
        sd[19:] = 1.0
        self.in_mean[19:] = 0.0
        self.in_std = np.where(sd > 1e-12, sd, 1.0).astype(np.float32)
        self.out_mean = Dt.mean(0).cpu().numpy().astype(np.float32)
        so = Dt.std(0).cpu().numpy().astype(np.float32)
        self.out_std = np.where(so > 1e-20, so, 1e-20).astype(np.float32)
        for name, w in (weights or {}).items():
            self.out_std[FEATURE_NAMES.index(name)] /= float(w)
        self.refresh()


# ---------------------------------------------------------------------------
# training data from the plain integrator
# ---------------------------------------------------------------------------
def reaction_targets(V, S, ct, dtype=torch.float64):
    """Increments of (V, 18 states) over N_SUB reaction-only substeps (V += -dt*I_ion),
    computed with the FP64 HIP ionic kernel."""
    n = V.shape[0]
    Vd = V.to(dtype).clone()
    st = S.to(dtype).t().contiguous()  # SoA (18, n), absolute values
    I = torch.empty(n, dtype=dtype, device=V.device)
    ctt = ct.to(torch.uint8).contiguous()
    s = nat.stream_handle(V.device)
    dt = T_MACRO / N_SUB
    for _ in range(N_SUB):
        nat.call("hp_ttp06_step_f64", Vd.data_ptr(), st.data_ptr(), I.data_ptr(), ctt.data_ptr(),
                 0, None, None, n, n, 1, dt, 0.0, 0.0, s)
        Vd -= dt * I
    return Vd, st.t()


def bigstep_states(V, S, ct):
    """One T_MACRO Rush-Larsen/forward-Euler step of the FP32 cell model, the
    arithmetic of hp_bigstep_reaction (state held relative to the Ki/Nai
    bases, as the FP32 engine stores it): returns (V + T_MACRO * (-I_ion),
    the advanced absolute state), FP64 tensors."""
    n = V.shape[0]
    st = S.float().clone()
    st[:, 0] -= KI_BASE
    st[:, 1] -= NAI_BASE
    st = st.t().contiguous()
    Vf = V.float().contiguous()
    I = torch.empty(n, dtype=torch.float32, device=V.device)
    ctt = ct.to(torch.uint8).contiguous()
    nat.call("hp_ttp06_step_f32", Vf.data_ptr(), st.data_ptr(), I.data_ptr(), ctt.data_ptr(), 0,
             None, None, n, n, 1, T_MACRO, KI_BASE, NAI_BASE, nat.stream_handle(V.device))
    S1 = st.t().double()
    S1[:, 0] += KI_BASE
    S1[:, 1] += NAI_BASE
    return Vf.double() - T_MACRO * I.double(), S1


def output_increments(V0, S0, V1, S1):
    """(n,19) increments in the network's output space (log features as log-ratios)."""
    D = torch.empty((V0.shape[0], N_OUT), dtype=torch.float64, device=V0.device)
    D[:, 0] = V1 - V0
    a, b = S0.double(), S1.double()
    lg = torch.from_numpy(IN_LOG[1:19].astype(bool)).to(V0.device)
    d = b - a
    d[:, lg] = torch.log(b[:, lg].clamp_min(1e-30)) - torch.log(a[:, lg].clamp_min(1e-30))
    D[:, 1:] = d
    return D.float()


def surrogate_mask(V, V_prev, dvdt_max=1.0, v_lo=-70.0, v_hi=-40.0):
    """True where the network (not the fallback) handles the voxel."""
    dvdt = (V - V_prev).abs() / T_MACRO
    return (dvdt <= dvdt_max) & ~((V > v_lo) & (V < v_hi))


def sample_from_engine(eng, n_macro, per_snapshot, seed=0, only_surrogate=True):
    """Run the plain engine and sample (V, state, phenotype) at macro-step
    boundaries; returns device tensors (V, S_abs, ct)."""
    g = torch.Generator(device=eng.dev).manual_seed(seed)
    sl = eng.slot_of_node.long()
    ct_all = (eng.label[sl] & 3)
    Vs, Ss, Cs = [], [], []
    V_prev = None
    for k in range(n_macro):
        eng.run(N_SUB, graph_chunk=N_SUB)
        V = eng.V[eng.p][sl].float()
        if V_prev is not None:
            cand = torch.arange(len(sl), device=eng.dev)
            if only_surrogate:
                cand = cand[surrogate_mask(V, V_prev)]
            if cand.numel():
                pick = cand[torch.randint(cand.numel(), (min(per_snapshot, cand.numel()),),
                                          generator=g, device=eng.dev)]
                st = eng.state[:, eng.state_of_node[pick].long()].t().float()
                st[:, 0] += eng.ki_base
                st[:, 1] += eng.nai_base
                Vs.append(V[pick])
                Ss.append(st)
                Cs.append(ct_all[pick])
        V_prev = V
    return torch.cat(Vs), torch.cat(Ss), torch.cat(Cs)


# ---------------------------------------------------------------------------
# training with the HIP kernels
# ---------------------------------------------------------------------------
class Trainer:
    """Mini-batch Adam on the MSE of normalised increments; every FLOP of the
    forward and backward pass runs in hrdaya HIP kernels."""

    def __init__(self, net: SurrogateNet, batch=8192, lr=2e-3, wd=0.0):
        self.net, self.B, self.lr, self.wd = net, batch, lr, wd
        dev = net.dev
        B = batch
        f32 = dict(dtype=torch.float32, device=dev)
        bf = dict(dtype=torch.bfloat16, device=dev)
        self.X0 = torch.zeros((B, 32), **bf)
        self.Z = [torch.empty((B, N_HID), **f32) for _ in range(3)]
        self.A = [torch.empty((B, N_HID), **bf) for _ in range(3)]
        self.Y = torch.empty((B, N_OUT), **f32)
        self.dY = torch.empty((B, N_OUT), **bf)
        self.dYf = torch.empty((B, N_OUT), **f32)
        self.dA = torch.empty((B, N_HID), **f32)
        self.dZ = torch.empty((B, N_HID), **bf)
        self.dZf = torch.empty((B, N_HID), **f32)
        self.gW = [torch.zeros_like(w) for w in net.W]
        self.gb = [torch.zeros_like(b) for b in net.b]
        self.mW = [torch.zeros_like(w) for w in net.W]
        self.vW = [torch.zeros_like(w) for w in net.W]
        self.mb = [torch.zeros_like(b) for b in net.b]
        self.vb = [torch.zeros_like(b) for b in net.b]
        self.loss = torch.zeros(1, **f32)
        self.step_no = 0

    def _gemm(self, M, N, K, A, sam, sak, B, sbk, sbn, Cout, ldc, bias=None, ksplit=1):
        nat.call("hp_gemm_bf16", M, N, K, A.data_ptr(), sam, sak, B.data_ptr(), sbk, sbn,
                 Cout.data_ptr(), ldc, 1.0, 0.0, nat.ptr(bias), ksplit, nat.stream_handle())

    def forward(self, Xn_bf16, B):
        """Xn_bf16: (B, 32) normalised inputs (first 22 columns used)."""
        net, s = self.net, nat.stream_handle()
        W = net.Wb
        self._gemm(B, N_HID, N_IN, Xn_bf16, 32, 1, W[0], 1, N_IN, self.Z[0], N_HID, net.b[0])
        for l in range(3):
            nat.call("hp_silu_bf16", self.Z[l].data_ptr(), self.A[l].data_ptr(), B, N_HID, N_HID,
                     N_HID, s)
            if l < 2:
                self._gemm(B, N_HID, N_HID, self.A[l], N_HID, 1, W[l + 1], 1, N_HID, self.Z[l + 1],
                           N_HID, net.b[l + 1])
        self._gemm(B, N_OUT, N_HID, self.A[2], N_HID, 1, W[3], 1, N_HID, self.Y, N_OUT, net.b[3])
        return self.Y[:B]

    def backward(self, Xn_bf16, Tn, B):
        """Gradients of the masked MSE (over the 19 outputs) into self.gW/self.gb."""
        net, s = self.net, nat.stream_handle()
        W = net.Wb
        self.loss.zero_()
        nat.call("hp_mse_grad", self.Y.data_ptr(), Tn.data_ptr(), self.dY.data_ptr(),
                 self.dYf.data_ptr(), B, N_OUT, N_OUT, self.loss.data_ptr(), s)
        ks = max(1, min(64, B // 1024))
        for g in self.gW:
            g.zero_()
        # output layer
        self._gemm(N_OUT, N_HID, B, self.dY, 1, N_OUT, self.A[2], N_HID, 1, self.gW[3], N_HID,
                   ksplit=ks)
        nat.call("hp_colsum", self.dYf.data_ptr(), B, N_OUT, N_OUT, self.gb[3].data_ptr(), s)
        self._gemm(B, N_HID, N_OUT, self.dY, N_OUT, 1, W[3], N_HID, 1, self.dA, N_HID)
        for l in (2, 1, 0):
            nat.call("hp_silu_bwd_bf16", self.Z[l].data_ptr(), self.dA.data_ptr(),
                     self.dZ.data_ptr(), self.dZf.data_ptr(), B * N_HID, s)
            nat.call("hp_colsum", self.dZf.data_ptr(), B, N_HID, N_HID, self.gb[l].data_ptr(), s)
            if l > 0:
                self._gemm(N_HID, N_HID, B, self.dZ, 1, N_HID, self.A[l - 1], N_HID, 1, self.gW[l],
                           N_HID, ksplit=ks)
                self._gemm(B, N_HID, N_HID, self.dZ, N_HID, 1, W[l], N_HID, 1, self.dA, N_HID)
            else:
                self._gemm(N_HID, N_IN, B, self.dZ, 1, N_HID, Xn_bf16, 32, 1, self.gW[0], N_IN,
                           ksplit=ks)

    def adam(self):
        self.step_no += 1
        s = nat.stream_handle()
        for p, g, m, v in list(zip(self.net.W, self.gW, self.mW, self.vW)) + list(
                zip(self.net.b, self.gb, self.mb, self.vb)):
            nat.call("hp_adam", p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), p.numel(),
                     self.lr, 0.9, 0.999, 1e-8, self.step_no, self.wd, s)
        for w, wb in zip(self.net.W, self.net.Wb):
            nat.call("hp_cast_bf16", w.data_ptr(), wb.data_ptr(), w.numel(), s)

    def prepare(self, X, D):
        """Normalised bf16 inputs (n,32) and fp32 targets (n,19) from raw features/increments."""
        net = self.net
        Xt = SurrogateNet.transform(X)
        Xn = torch.zeros((X.shape[0], 32), dtype=torch.bfloat16, device=X.device)
        Xn[:, :N_IN] = ((Xt - torch.from_numpy(net.in_mean).to(X.device))
                        / torch.from_numpy(net.in_std).to(X.device)).to(torch.bfloat16)
        Tn = ((D - torch.from_numpy(net.out_mean).to(X.device))
              / torch.from_numpy(net.out_std).to(X.device)).float().contiguous()
        return Xn, Tn

    def fit(self, Xn, Tn, epochs=1, lr_decay=0.5, log=None, seed=0):
        n = Xn.shape[0]
        B = self.B
        g = torch.Generator(device=Xn.device).manual_seed(seed)
        hist = []
        t0 = time.perf_counter()
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g, device=Xn.device)
            tot, cnt = 0.0, 0
            for i in range(0, n - B + 1, B):
                idx = perm[i:i + B]
                xb = Xn[idx].contiguous()
                tb = Tn[idx].contiguous()
                self.forward(xb, B)
                self.backward(xb, tb, B)
                self.adam()
                if cnt % 50 == 0:
                    tot += float(self.loss)
                cnt += 1
            hist.append(tot / max(1, (cnt + 49) // 50))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.3e} ({time.perf_counter() - t0:.1f} s)")
            self.lr *= lr_decay
        return hist


# ---------------------------------------------------------------------------
# surrogate form of the time loop
# ---------------------------------------------------------------------------
class SurrogateRunner:
    """Drives a MonodomainEngine (FP32) in a multirate form: per macro-step one
    reaction update for the non-fallback voxels, then N_SUB substeps in which
    the fallback voxels run the full TTP06 step.

    mode "mlp": the learned network (hp_surrogate_step; a correcting network,
    SurrogateNet.correct, runs after hp_bigstep_reaction in the same call);
    mode "bigstep": one 0.2 ms Rush-Larsen step of the exact model
    (hp_bigstep_reaction) -- the same rule without the network.

    compact=True: the reaction call compacts the fallback voxels into a list;
    each substep runs hp_fallback_react over that list, then the diffusion-only
    fused kernel (do_reaction = 0, dV/dt from the rate array).  compact=False:
    each substep runs the fused kernel with do_reaction = 2 (fallback voxels
    found by their NaN rate, in divergent waves).  Both give identical results.
    """

    def __init__(self, eng, net: SurrogateNet = None, mode="mlp", dvdt_max=1.0, v_lo=-70.0,
                 v_hi=-40.0, n_blocks=0, compact=True, graph_macros=2, variant=2,
                 rest_v=REST_V, rest_rate=REST_RATE, corr_max=CORR_MAX):
        if eng.dtype != torch.float32:
            raise ValueError("the surrogate form runs on the FP32 engine")
        # substeps: the engine's own diffusion scheme, the reaction from the rate
        # array (CN: the reaction pass runs the cell model for the NaN-rate
        # fallback voxels itself; explicit: compacted fallback list + the
        # diffusion-only fused kernel)
        if mode not in ("mlp", "bigstep"):
            raise ValueError(mode)
        if mode == "mlp" and net is None:
            raise ValueError("mode 'mlp' needs a trained SurrogateNet")
        self.eng, self.net, self.mode = eng, net, mode
        self.rate = torch.full((eng.lay.n_slots,), float("nan"), dtype=torch.float32,
                               device=eng.dev)
        self.V_prev = torch.full((eng.lay.n_slots,), float("nan"), dtype=torch.float32,
                                 device=eng.dev)
        self.compact = bool(compact) and eng.scheme == "explicit"
        self.fb_list = torch.zeros(eng.lay.n_slots if self.compact else 1, dtype=torch.int32,
                                   device=eng.dev)
        self.fb_count = torch.zeros(1, dtype=torch.int32, device=eng.dev)
        # the inference kernel visits the tissue slots only (padding of partly
        # filled bricks skipped), listed in slot order, padded to 32 with -1
        tissue = torch.nonzero(eng.label != 0xFF).flatten().to(torch.int32)
        n_list = (tissue.numel() + 31) // 32 * 32
        self.slot_list = torch.full((max(32, n_list),), -1, dtype=torch.int32, device=eng.dev)
        self.slot_list[:tissue.numel()] = tissue
        # tissue slots in slot order: their compact state indices are 0, 1, ...
        self.state_list = torch.zeros_like(self.slot_list)
        self.state_list[:tissue.numel()] = torch.arange(tissue.numel(), dtype=torch.int32,
                                                        device=eng.dev)
        # inference kernel shape: two 16-voxel MFMA columns per wave; variant 2
        # (1024 threads, four waves per SIMD) 323 us per call on the 2 M-voxel
        # BiV against 368 us for variant 1 (512 threads)
        self.sp = nat.HpSurrogateParams(T_MACRO, dvdt_max, v_lo, v_hi, eng.ki_base, eng.nai_base,
                                        0, int(variant),
                                        nat.ptr(self.fb_list) if self.compact else None,
                                        nat.ptr(self.fb_count) if self.compact else None,
                                        nat.ptr(self.slot_list), self.slot_list.numel(),
                                        nat.ptr(self.state_list),
                                        int(mode == "mlp" and net.correct), rest_v, rest_rate,
                                        corr_max)
        self.n_blocks = n_blocks
        self._orig = (eng.p_myo.do_reaction, eng.p_myo.rate_ext)
        self._graphs = {}
        self.graph_macros = int(graph_macros)

    def launch_reaction(self):
        eng = self.eng
        if self.compact:
            self.fb_count.zero_()  # a fill kernel (safe inside a captured graph)
        if self.mode == "mlp" and self.net.correct:
            # the big step decides, lists the fallback voxels and steps the
            # rest; the network then adds its correction
            nat.call("hp_bigstep_reaction", C.byref(eng.bricks), nat.ptr(eng.V[eng.p]),
                     nat.ptr(self.V_prev), nat.ptr(eng.state), eng.state_ld,
                     nat.ptr(self.rate), C.byref(self.sp), nat.stream_handle())
        if self.mode == "mlp":
            nat.call("hp_surrogate_step", C.byref(eng.bricks), nat.ptr(eng.V[eng.p]),
                     nat.ptr(self.V_prev), nat.ptr(eng.state), eng.state_ld,
                     nat.ptr(self.rate), C.byref(self.net.hp_w), C.byref(self.net.hp_norm),
                     C.byref(self.sp), self.n_blocks, nat.stream_handle())
        else:
            nat.call("hp_bigstep_reaction", C.byref(eng.bricks), nat.ptr(eng.V[eng.p]),
                     nat.ptr(self.V_prev), nat.ptr(eng.state), eng.state_ld,
                     nat.ptr(self.rate), C.byref(self.sp), nat.stream_handle())

    def launch_substep(self):
        """One substep: fallback voxels' TTP06 step, then diffusion + rate."""
        eng = self.eng
        if self.compact:
            nat.call("hp_fallback_react_f32", C.byref(eng.bricks), nat.ptr(eng.V[eng.p]),
                     nat.ptr(eng.state), eng.state_ld, nat.ptr(self.rate),
                     nat.ptr(self.fb_list), nat.ptr(self.fb_count), eng.lay.n_slots,
                     C.byref(eng.p_myo), nat.stream_handle())
        eng.p_myo.do_reaction = 0 if self.compact else 2  # CN engines: always 2
        eng.p_myo.rate_ext = nat.ptr(self.rate)
        try:
            eng._launch_step()
        finally:
            eng.p_myo.do_reaction, eng.p_myo.rate_ext = self._orig

    def _launch_macro(self):
        self.launch_reaction()
        for _ in range(N_SUB):
            self.launch_substep()

    def run(self, n_macro, graph=True, adaptive=False):
        """Advance n_macro macro-steps (graphs of graph_macros at a time).

        adaptive (CN engines with Jacobi sweeps): the substeps' sweep bound is
        fitted per graph as MonodomainEngine.run(adaptive=True) fits it, the
        graph replayed from its kept state with the full bound if a substep
        stopped at the fitted one (the runner's rate, V_prev and fallback
        lists are kept with the engine's state)."""
        if not graph:
            for _ in range(n_macro):
                self._launch_macro()
            return
        eng = self.eng
        if adaptive and eng.cn is not None and eng.cn_solver == "jacobi":
            k = self.graph_macros
            extra = [self.rate, self.V_prev, self.fb_list, self.fb_count]
            eng._run_chunks_adaptive(n_macro // k, lambda: self._graph().replay(), extra)
            for _ in range(n_macro % k):
                self._launch_macro()
            return
        g = self._graph()
        k = self.graph_macros
        for _ in range(n_macro // k):
            g.replay()
        for _ in range(n_macro % k):
            self._launch_macro()

    def _graph(self):
        eng = self.eng
        key = (eng.p, eng.fuse_cable, eng.overlap_cable,  # buffers as captured
               eng.cn.max_sweeps if eng.cn is not None else 0)
        if self._graphs.get(key) is None:
            p0 = eng.p
            g = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream(device=eng.dev)
            side.wait_stream(torch.cuda.current_stream(eng.dev))
            k = self.graph_macros
            if (k * N_SUB) % 2:
                raise ValueError("a captured graph must hold an even number of substeps")
            with torch.cuda.stream(side):
                with torch.cuda.graph(g, stream=side):
                    for _ in range(k):
                        self._launch_macro()
            torch.cuda.current_stream(eng.dev).wait_stream(side)
            eng.p = p0
            self._graphs[key] = g
        return self._graphs[key]

    def _fallback_mask(self):
        """(n_slots,) bool: fallback voxels of the current macro-step."""
        if not self.compact:
            return torch.isnan(self.rate) & (self.eng.label != 0xFF)
        m = torch.zeros(self.eng.lay.n_slots, dtype=torch.bool, device=self.eng.dev)
        n = int(self.fb_count.item())
        m[self.fb_list[:n].long()] = True
        return m

    def fallback_fraction(self):
        live = self.eng.label != 0xFF
        return float(self._fallback_mask()[live].float().mean())

    def fallback_brick_fraction(self):
        """Fraction of bricks holding a fallback voxel (the waves that would
        diverge without compaction)."""
        fb = self._fallback_mask().view(-1, self.eng.lay.bvol)
        return float(fb.any(dim=1).float().mean())


def training_slab(device="cuda", nx=64, ny=64, nz=16, h=0.25, seed=0, onset=0.0):
    """The 64x64x16 @ 0.25 mm slab (sigma 0.13 isotropic, corner stimulus
    52 pA/pF for 2 ms at `onset` ms) with a transmural phi across z, so that
    endo, mid and epi cells are all sampled.  Returns an FP32 MonodomainEngine."""
    from .anisotropic_simulation.functions.purkinje_myocardium_anisotropic_solver import (
        build_coupled_system)
    from .geometry import slab, stub_tree
    S, phi, f = slab(nx, ny, nz)
    phi = (np.arange(nz)[None, None, :] / max(1, nz - 1) * np.ones((nx, ny, 1))).astype(np.float32)
    region = np.zeros(S.shape, bool)
    region[:4, :4, :4] = True
    nodes, el, act = stub_tree(S.shape)
    eng, _ = build_coupled_system(
        nodes, el, act, S, phi, f, voxel_size=h, dx_p=h, sigma_P=0.0, sigma_l=0.13, sigma_t=0.13,
        Cm=0.01, A_M=140.0, c_pmj=0.0, T=500.0, phi_endo_max=1 / 3, phi_epi_min=2 / 3,
        ectopic_region=region, dt=0.02, dtype=torch.float32, device=device,
        stim_protocol=[dict(name="corner", target="ectopic", amp=0.52, dur=2.0, onset=onset)])
    return eng


def training_biv(device="cuda", n=160, seed=1, T=600.0):
    """A BiV shell of another seed and size than the 2 M-voxel benchmark one
    (n^3 grid, voxel 0.4*256/n mm; Laplace transmural coordinate and helix
    fibres; endo/mid/epi thirds; a seeded Purkinje tree coupled through 800
    PMJs; root and apical stimuli), FP32 engine with the CN step solved by
    Jacobi sweeps.  Returns the engine."""
    from . import geometry as G
    from .anisotropic_simulation.functions.purkinje_myocardium_anisotropic_solver import (
        build_coupled_system)
    h = 0.4 * 256 / n
    geo = G.biv_shell(n, h, seed=seed, coords="laplace", device=device)
    nodes, el, act = G.purkinje_tree(geo, 800, seed=seed)
    eng, _ = build_coupled_system(
        nodes, el, act, geo["S"], geo["phi"], geo["fibre_dir"], voxel_size=h, dx_p=0.2,
        sigma_P=0.45, sigma_l=0.24, sigma_t=0.08, Cm=0.01, A_M=140.0, R_P=0.015, c_pmj=2.0,
        n_pmj=1, stim_len_mm=2.0, T=T, phi_endo_max=1.0 / 3.0, phi_epi_min=2.0 / 3.0,
        ectopic_region=G.apical_stimulus_region(geo, 3.0), dt=0.02, dtype=torch.float32,
        device=device, scheme="cn", cn_solver="jacobi", max_sweeps=5,
        stim_protocol=[dict(name="av_node", target="root", amp=12.0, dur=2.0, onset=0.0),
                       dict(name="apex", target="ectopic", amp=0.52, dur=2.0, onset=0.0)])
    return eng


def training_runs(device="cuda", samples=20_000_000, runs=((0.0, 500.0), (150.0, 500.0)), seed=0,
                  correct=True, biv=None):
    """Training set from plain-integrator runs of the slab: (stimulus onset, length)
    per run.  The onset-0 run is the reference's slab protocol; the delayed
    onset adds cells resting from the initial state for longer than the slab's
    activation takes, which tissue larger than the slab contains.  biv = (n,
    seed): one more run, a 600 ms beat of a BiV shell (training_biv) with its
    Purkinje-driven activation and anisotropic fronts.  Returns (X raw
    features, D targets) on the device."""
    Xs, Ds = [], []
    plan = [("slab", onset, T) for onset, T in runs]
    if biv is not None:
        plan.append(("biv", biv, 600.0))
    per_run = samples // len(plan)
    for k, (kind, arg, T) in enumerate(plan):
        eng = training_slab(device, onset=arg) if kind == "slab" else training_biv(
            device, n=arg[0], seed=arg[1], T=T)
        n_macro = int(round(T / T_MACRO))
        X, D = build_training_set(eng, n_macro, max(1, per_run // n_macro), seed=seed + k,
                                  correct=correct)
        Xs.append(X)
        Ds.append(D)
        del eng
    return torch.cat(Xs), torch.cat(Ds)


def build_training_set(eng, n_macro, per_snapshot, seed=0, correct=True):
    """Samples of the plain integrator's trajectories and their 10-substep
    reaction increments: returns (X raw features, D targets).  correct=True:
    the features hold the state after the big step and the targets are what
    the ten substeps add to the big step's result; False: the state at the
    start and the whole increment."""
    V, St, ct = sample_from_engine(eng, n_macro, per_snapshot, seed=seed)
    V1, S1 = reaction_targets(V, St, ct)
    if correct:  # the voxels the network corrects: not at rest (REST_V, REST_RATE)
        Vb, Sb = bigstep_states(V, St, ct)
        keep = ~((V.double() < REST_V) & ((Vb - V.double()).abs() / T_MACRO < REST_RATE))
        return (SurrogateNet.features(V[keep], Sb[keep].float(), ct[keep]),
                output_increments(Vb[keep], Sb[keep], V1[keep], S1[keep]))
    D = output_increments(V.double(), St.double(), V1, S1)
    return SurrogateNet.features(V, St, ct), D


def train_surrogate(X, D, device="cuda", epochs=14, batch=8192, lr=2e-3, lr_decay=0.75,
                    holdout=0.05, log=None, seed=0, weights=None, correct=True):
    """Fit a SurrogateNet with the HIP training kernels; returns (net, trainer, stats).
    ``correct`` says which targets (X, D) hold (build_training_set); the
    held-out dV error is that of the network's macro-step, and in correction
    mode the big step's own error on the same samples is reported beside it
    (heldout_bigstep_dv_mae/max)."""
    n = X.shape[0]
    nte = max(batch, int(holdout * n))
    ntr = n - nte
    # a random held-out set (the samples come in the order of the runs' time)
    perm = torch.randperm(n, generator=torch.Generator(device=X.device).manual_seed(seed + 7),
                          device=X.device)
    X, D = X[perm], D[perm]
    del perm
    net = SurrogateNet(device, seed=seed)
    net.correct = bool(correct)
    net.fit_normalisation(SurrogateNet.transform(X[:ntr]), D[:ntr],
                          LOSS_WEIGHTS if weights is None else weights)
    tr = Trainer(net, batch=batch, lr=lr)
    Xn, Tn = tr.prepare(X, D)
    t0 = time.perf_counter()
    hist = tr.fit(Xn[:ntr], Tn[:ntr], epochs=epochs, lr_decay=lr_decay, log=log, seed=seed)
    net.refresh_inference()
    train_s = time.perf_counter() - t0
    # held-out error of dV over the macro-step (mV)
    preds = []
    for i in range(ntr, n - batch + 1, batch):
        preds.append(tr.forward(Xn[i:i + batch].contiguous(), batch).clone())
    m = len(preds) * batch
    Yp = torch.cat(preds) * torch.from_numpy(net.out_std).to(X.device) + torch.from_numpy(
        net.out_mean).to(X.device)
    err = (Yp[:, 0] - D[ntr:ntr + m, 0]).abs()
    st = dict(loss=hist, train_s=train_s, n_train=ntr, heldout_dv_mae=float(err.mean()),
              heldout_dv_max=float(err.max()), correct=bool(correct))
    if correct:  # the target is the big step's error: its size is the big step's own error
        bs = D[ntr:ntr + m, 0].abs()
        st.update(heldout_bigstep_dv_mae=float(bs.mean()), heldout_bigstep_dv_max=float(bs.max()))
    return net, tr, st
