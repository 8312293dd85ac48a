"""Write golden per-step TTP06 vectors by running the reference's own module.

Run in the development container only (the reference tree is not present on
GPU machines); the output ``ttp06_golden.npz`` is committed next to this
script and is what the tests read.

What is imported: ``hrdayapy/simulation/functions/ionic_models/gpu_ttp06.py``
and the constants module it imports (``ttp06_phenotype_constants.py``), loaded
by file path under a synthetic parent package so that the module's relative
import resolves to the real constants file. Nothing else of the reference is
loaded (the package ``__init__`` chain needs pynrrd/pyvista, which are absent).

Why an import: per-step, formula-level parity of ``_currents_and_gates``
(gpu_ttp06.py:198-412) is the one hot-path function that no closed form or
literature value pins at the level of a single step.

Vectors:
  * ``V`` (N,), ``state`` (N,18): seeded random inputs spanning rest, upstroke,
    plateau and repolarisation (V in [-95, 55] mV, gates in (0,1),
    concentrations around their physiological ranges), plus the exact resting
    state and the I_CaL removable point V = 15 mV;
  * for each phenotype p in (endo, mid, epi): ``I_<p>`` (N,) and ``S_<p>``
    (N,18) after one step of dt = 0.02 ms;
  * ``trace_epi``: the single-cell epicardial AP of experiment 1's protocol
    (52 pA/pF for 1 ms, dt 0.02 ms), first 500 ms, every 10th step.
"""
from __future__ import annotations

import importlib.util
import sys
import types
from pathlib import Path

import numpy as np
import torch

REF = Path("/root/reference/hrdayapy/simulation/functions/ionic_models")
OUT = Path(__file__).resolve().parent / "ttp06_golden.npz"


def _load_reference():
    pkg = types.ModuleType("_ref_ionic")
    pkg.__path__ = [str(REF)]
    sys.modules["_ref_ionic"] = pkg
    for name in ("ttp06_phenotype_constants", "gpu_ttp06"):
        spec = importlib.util.spec_from_file_location(f"_ref_ionic.{name}", REF / f"{name}.py")
        mod = importlib.util.module_from_spec(spec)
        sys.modules[spec.name] = mod
        spec.loader.exec_module(mod)
    return sys.modules["_ref_ionic.gpu_ttp06"]


def _inputs(n=1024, seed=20260101):
    rng = np.random.default_rng(seed)
    V = rng.uniform(-95.0, 55.0, n)
    st = np.empty((n, 18))
    st[:, 0] = rng.uniform(130.0, 145.0, n)      # Ki
    st[:, 1] = rng.uniform(6.0, 14.0, n)         # Nai
    st[:, 2] = 10 ** rng.uniform(-4.5, -2.8, n)  # Cai
    for c in range(3, 15):                       # gates
        st[:, c] = rng.uniform(1e-6, 1.0, n)
    st[:, 15] = rng.uniform(0.5, 5.0, n)         # CaSR
    st[:, 16] = rng.uniform(0.3, 1.0, n)         # Rprime
    st[:, 17] = 10 ** rng.uniform(-4.5, -1.5, n) # Cass
    ref = _load_reference()
    init = ref._init_tensor(1, torch.device("cpu"), torch.float64).numpy()[0]
    st[0] = init
    V[0] = -86.709
    st[1] = init
    V[1] = 15.0
    V[2] = 15.0 + 1e-9
    V[3] = -40.0
    return V, st


def main():
    ref = _load_reference()
    V, st = _inputs()
    out = {"V": V, "state": st, "dt": np.float64(0.02)}
    classes = {"endo": ref.EndoTTP06, "mid": ref.MidTTP06, "epi": ref.EpiTTP06}
    for name, cls in classes.items():
        I, S = cls().step(torch.from_numpy(V), torch.from_numpy(st), 0.02, 0.0)
        out[f"I_{name}"] = I.numpy()
        out[f"S_{name}"] = S.numpy()
    # single-cell AP (experiment 1 protocol), epi phenotype
    model = ref.EpiTTP06()
    s = model.init_state_tensor(1, torch.device("cpu"), torch.float64)
    v = torch.full((1,), -86.709, dtype=torch.float64)
    n_steps, stim_steps = 25000, 50
    trace = np.empty(n_steps // 10)
    for k in range(n_steps):
        i_st = 52.0 if k < stim_steps else 0.0
        I, s = model.step(v, s, 0.02, i_st)
        v = v + 0.02 * (i_st - I)
        if k % 10 == 0:
            trace[k // 10] = float(v[0])
    out["trace_epi"] = trace
    out["trace_epi_state_end"] = s.numpy()[0]
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
