"""The allocation-free residual entry point (hp_poisson_residual_f64_noalloc)
against hp_poisson_residual_f64, through the C ABI: the same r and the same
||r||^2, ||b||^2 bit for bit, on two calls in a row, the second on a smaller
torso grid (the scratch array keeps the first call's partials)."""
import ctypes as C

import numpy as np
import pytest
import torch

from hrdayapy_amd import _native as nat
from hrdayapy_amd.ecg.functions.forward import PoissonSolver
from hrdayapy_amd.ecg.functions.torso import TorsoGrid
from tests.test_oracle_ecg import _torso

pytestmark = pytest.mark.gpu


def _case(dev, n, seed):
    lab, smap = _torso(n=n)
    g = TorsoGrid.from_labels(lab, smap, 1.5)
    ps = PoissonSolver(g, dev)
    rng = np.random.default_rng(seed)
    x = torch.from_numpy(rng.standard_normal(ps.n)).to(dev)
    b = torch.from_numpy(rng.standard_normal(ps.n)).to(dev)
    return ps, b, x


def _both(ps, b, x):
    s = nat.stream_handle()
    r_ws = torch.full_like(x, float("nan"))
    nat.call("hp_poisson_residual_f64", C.byref(ps.torso), b.data_ptr(), x.data_ptr(),
             r_ws.data_ptr(), C.byref(ps.ws), s)
    out_ws = ps.out[:2].clone()
    r_na = torch.full_like(x, float("nan"))
    out_na = torch.full((2,), float("nan"), dtype=torch.float64, device=x.device)
    nat.call("hp_poisson_residual_f64_noalloc", C.byref(ps.torso), b.data_ptr(), x.data_ptr(),
             r_na.data_ptr(), out_na.data_ptr(), s)
    torch.cuda.synchronize()
    return r_ws.cpu().numpy(), out_ws.cpu().numpy(), r_na.cpu().numpy(), out_na.cpu().numpy()


def test_noalloc_residual_matches_workspace_entry_point(cuda):
    # the larger grid first (many blocks of partials), then a smaller one
    for n, seed in (((72, 64, 160), 1), ((30, 26, 22), 2)):
        print(f"[case] residual noalloc torso {n} seed {seed}", flush=True)
        ps, b, x = _case(cuda, n, seed)
        r_ws, o_ws, r_na, o_na = _both(ps, b, x)
        assert np.isfinite(r_na).all() and np.isfinite(o_na).all()
        assert np.array_equal(r_ws, r_na)
        assert np.array_equal(o_ws, o_na)
        assert o_na[0] > 0 and o_na[1] > 0


def test_noalloc_residual_rejects_missing_output(cuda):
    ps, b, x = _case(cuda, (30, 26, 22), 3)
    lib = nat.load()
    rc = lib.hp_poisson_residual_f64_noalloc(C.byref(ps.torso), b.data_ptr(), x.data_ptr(),
                                             x.data_ptr(), None, nat.stream_handle())
    assert rc != 0 and "required" in nat.last_error()
