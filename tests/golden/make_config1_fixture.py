"""Oracle fixture of configuration 1 over its activation window.

Configuration 1: the 64x64x16 all-myocardium slab at 0.25 mm, isotropic
0.13 S/m, endocardial TTP06 (phi = 0), FP64, dt = 0.02 ms, 52 pA/pF
(0.52 uA/mm^2 at Cm = 0.01) for 2 ms on the 4x4x4 corner block, inert
Purkinje stub.  The oracle (the reference's scheme: Lie splitting, CN with
BiCGSTAB to 1e-5) runs the first 50 ms -- the wave crosses the whole slab --
and this script stores the activation time of every voxel and Vm of eight
probe voxels every 0.5 ms (config1_oracle.npz, read by
tests/test_gpu_configs.py).  The oracle is slow on a CPU (the reference's
eager TTP06 step), hence the stored run.  Usage: python make_config1_fixture.py
"""
from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np
import torch

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from hrdayapy_amd import geometry as G  # noqa: E402
from oracle import monodomain as OMD  # noqa: E402

T_MS = 50.0
PROBES = np.array([[1, 1, 1], [10, 10, 5], [20, 40, 8], [32, 32, 8], [40, 10, 2], [50, 50, 12],
                   [63, 63, 15], [63, 0, 0]])


def config1_problem(T=T_MS):
    S, _, fib = G.slab()
    nodes, el, act = G.stub_tree(S.shape)
    region = np.zeros(S.shape, dtype=bool)
    region[:4, :4, :4] = True
    kw = dict(voxel_size=0.25, dx_p=0.25, sigma_P=0.0, sigma_l=0.13, sigma_t=0.13, Cm=0.01,
              A_M=140.0, c_pmj=0.0, T=T, ectopic_region=region, phi_endo_max=0.35,
              phi_epi_min=0.65,
              stim_protocol=[dict(name="corner", target="ectopic", amp=0.52, dur=2.0, onset=0.0)])
    return (nodes, el, act, S, np.zeros(S.shape, np.float32), fib), kw


def main():
    torch.set_num_threads(int(sys.argv[1]) if len(sys.argv) > 1 else 4)
    args, kw = config1_problem()
    prob = OMD.Problem(*args, **kw)
    S = args[3]
    flat = np.ravel_multi_index(tuple(PROBES.T), S.shape)
    node = np.searchsorted(np.ravel_multi_index(tuple(prob.myo["vox_idx"].T), S.shape), flat)
    t0 = time.time()
    r = OMD.run(prob, T_MS, 0.02, scheme="cn", record_every=25,
                record_nodes=prob.Np + node)
    print(f"oracle {r['steps']} steps in {time.time() - t0:.0f} s")
    np.savez_compressed(HERE / "config1_oracle.npz", at=r["at"][prob.Np:].astype(np.float64),
                        probe_vox=PROBES, probe_V=r["trace_V"], probe_t=r["trace_t"],
                        T_ms=T_MS)


if __name__ == "__main__":
    main()
