"""Batched ensembles: many independent tissue instances in one brick list.

The reference runs one tissue per call, with the ionic conductances as
class-level scalars (gpu_ttp06.py:428-429) and sigma_l/sigma_t as call
arguments (purkinje_myocardium_anisotropic_solver.py:1321-1322).  Here the
instances of an ensemble -- slabs and biventricular shells of different
sizes, each with its own conductivity scaling, g_Ks/g_to scaling and pacing
sites -- are laid out back to back in ONE brick list: every time step is one
launch of the fused kernel over the bricks of all instances, so the uneven
instance sizes cost nothing in load balance (a brick is a brick), and K steps
are replayed from one captured HIP graph.

Per-instance data on the device:
  * inst id per voxel slot (uint16), selecting a row of inst_params
    (g_Ks scale, g_to scale) inside the cell step;
  * D6 per voxel, pre-multiplied by the instance's conductivity scales and by
    (h_ref/h)^2, so that instances of different voxel size share one
    diffusion scale 1/(A_M h_ref^2) (the 19-point weights are homogeneous of
    degree one in D);
  * stimulus groups (inst << 3) | local site combination, evaluated on the
    device from the per-instance site table.
Instances never exchange current: a brick's neighbour table only names
bricks of its own instance.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import geometry as G
from . import tissue
from .engine import MonodomainEngine
from .layout import BrickLayout, choose_brick

H_REF = 0.4  # mm; the common diffusion scale of an ensemble engine
BIV_VOXELS_256 = 2_006_132  # voxels of biv_shell(256, 0.4)


@dataclass
class InstanceSpec:
    kind: str                 # "slab" or "biv"
    n_target: int             # requested voxel count
    seed: int
    sigma_l: float            # mS/mm, already scaled
    sigma_t: float
    gks_scale: float = 1.0
    gto_scale: float = 1.0
    sites: list = field(default_factory=list)  # dict(pick in [0,1), radius_mm, onset, dur, amp)


def draw_instances(n_inst=64, seed=0, size_range=(200_000, 2_000_000), kinds=("slab", "biv"),
                   sigma_l=0.24, sigma_t=0.08, sigma_jitter=0.3, g_jitter=0.5, max_sites=3,
                   onset_range=(0.0, 20.0), amp=0.52, dur=2.0, radius_mm=1.0):
    """Seeded ensemble description: sizes uniform in size_range, sigma_l and
    sigma_t each scaled by U(1-sigma_jitter, 1+sigma_jitter), g_Ks and g_to by
    U(1-g_jitter, 1+g_jitter), 1..max_sites pacing sites at random tissue
    voxels with onsets uniform in onset_range."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_inst):
        kind = kinds[int(rng.integers(len(kinds)))]
        n = int(rng.uniform(*size_range))
        sl = sigma_l * rng.uniform(1 - sigma_jitter, 1 + sigma_jitter)
        st = sigma_t * rng.uniform(1 - sigma_jitter, 1 + sigma_jitter)
        gks = rng.uniform(1 - g_jitter, 1 + g_jitter)
        gto = rng.uniform(1 - g_jitter, 1 + g_jitter)
        sites = [dict(pick=rng.uniform(0, 1), onset=float(rng.uniform(*onset_range)), dur=dur,
                      amp=amp, radius_mm=radius_mm)
                 for _ in range(int(rng.integers(1, max_sites + 1)))]
        out.append(InstanceSpec(kind, n, int(rng.integers(1 << 31)), float(sl), float(st),
                                float(gks), float(gto), sites))
    return out


def instance_geometry(spec: InstanceSpec):
    """(S, phi, fibre_dir, h) of one instance, sized to about spec.n_target voxels.

    slab: 0.25 mm voxels, 16 voxels thick, fibres along x, phi linear across
    the thickness (endo/mid/epi thirds); biv: biv_shell at 0.4 mm on an n^3
    grid with n chosen from the voxel count of the 256^3 shell.
    """
    if spec.kind == "slab":
        nz = 16
        side = max(8, int(round(np.sqrt(spec.n_target / nz))))
        S, _, f = G.slab(side, side, nz)
        phi = np.broadcast_to((np.arange(nz) / (nz - 1)).astype(np.float32),
                              S.shape).copy()
        return S, phi, f, 0.25
    if spec.kind == "biv":
        n = int(round(256 * (spec.n_target / BIV_VOXELS_256) ** (1.0 / 3.0)))
        n += n % 2
        geo = G.biv_shell(max(n, 32), H_REF, seed=spec.seed)
        return geo["S"], geo["phi"], geo["fibre_dir"], H_REF
    raise ValueError(f"unknown instance kind {spec.kind!r}")


def instance_geometries(specs, workers=8):
    """instance_geometry for every spec, on a pool of forked worker processes
    (host numpy/scipy work only; the workers never touch the GPU)."""
    specs = list(specs)
    if workers <= 1 or len(specs) <= 1:
        return [instance_geometry(s) for s in specs]
    import multiprocessing as mp
    with mp.get_context("fork").Pool(min(workers, len(specs))) as pool:
        return pool.map(instance_geometry, specs, chunksize=1)


def _site_nodes(spec_site, vox, h, rng_pick):
    """Myocardial node ids within radius of a tissue voxel chosen by rng_pick in [0,1)."""
    centre = vox[min(len(vox) - 1, int(rng_pick * len(vox)))].astype(np.float64)
    r = spec_site["radius_mm"] / h
    d2 = ((vox - centre) ** 2).sum(axis=1)
    return np.nonzero(d2 <= r * r)[0]


class Ensemble:
    """Instances concatenated into one MonodomainEngine (no Purkinje network).

    ``node_offset[i]`` .. ``node_offset[i+1]`` are instance i's nodes in the
    engine's node order (each instance's own np.argwhere(S) order).
    """

    def __init__(self, specs, *, geoms=None, dt=0.02, dtype=torch.float32, device="cuda",
                 brick=None, A_M=140.0, Cm=0.01, phi_endo_max=1 / 3, phi_epi_min=2 / 3,
                 scheme="explicit", theta=0.5, cg_tol=1e-5, max_sweeps=10):
        """scheme "cn": the reference's Crank-Nicolson step for every instance,
        solved by Jacobi sweeps (a Krylov solve of the batch would couple the
        instances through its inner products; the sweeps never do).
        The reference solves one instance per call, each to ||r|| <= cg_tol
        ||b|| of its own system; the batch is stopped on its whole residual, so
        the batch tolerance is cg_tol * sqrt(smallest instance / all voxels),
        which bounds every instance's own relative residual by about cg_tol
        (its ||b|| scales with the square root of its voxel count at the
        voltages of a beat)."""
        t0 = time.perf_counter()
        self.specs = list(specs)
        geoms = geoms if geoms is not None else instance_geometries(self.specs)
        self.t_geometry = time.perf_counter() - t0
        if brick is None:
            brick = choose_brick(sum(int(np.count_nonzero(g[0])) for g in geoms))
        lays, D6s, cts, insts, sites = [], [], [], [], []
        off = 0
        self.h = []
        for i, (spec, (S, phi, f, h)) in enumerate(zip(self.specs, geoms)):
            lay = BrickLayout(S, brick)
            vox = lay.vox_idx
            D6 = tissue.conductivity_D6(f, vox, spec.sigma_l, spec.sigma_t)
            D6 *= (H_REF / h) ** 2
            lays.append(lay)
            D6s.append(D6)
            cts.append(tissue.phenotype_from_phi(phi, vox, phi_endo_max, phi_epi_min))
            insts.append(np.full(lay.n_nodes, i, dtype=np.uint16))
            for s in spec.sites:
                ids = _site_nodes(s, vox.astype(np.float64), h, s["pick"])
                sites.append(dict(node_ids=off + ids, amp=s["amp"], dur=s["dur"],
                                  onset=s["onset"], period=None, n_pulses=None, inst=i))
            self.h.append(h)
            off += lay.n_nodes
        self.layouts = lays
        self.layout = BrickLayout.concat(lays)
        self.node_offset = self.layout.node_offset
        inst_params = np.zeros((len(self.specs), 4))
        inst_params[:, 0] = [s.gks_scale for s in self.specs]
        inst_params[:, 1] = [s.gto_scale for s in self.specs]
        self.sites = sites
        sizes = np.diff(self.node_offset)
        self.batch_tol = float(cg_tol * np.sqrt(sizes.min() / max(1, sizes.sum())))
        self.eng = MonodomainEngine(
            self.layout, np.concatenate(D6s, axis=1), np.concatenate(cts), h=H_REF, A_M=A_M,
            Cm=Cm, dt=dt, dtype=dtype, device=device, stim_sites=sites,
            inst=np.concatenate(insts), inst_params=inst_params, scheme=scheme, theta=theta,
            cg_tol=self.batch_tol, max_sweeps=max_sweeps, cn_solver="jacobi")
        self.t_setup = time.perf_counter() - t0

    @property
    def n_nodes(self):
        return self.layout.n_nodes

    def run(self, n_steps, graph_chunk=50):
        self.eng.run(n_steps, graph_chunk=graph_chunk)

    def split(self, x):
        """Split a per-node array (engine node order) into per-instance pieces."""
        o = self.node_offset
        return [x[o[i]:o[i + 1]] for i in range(len(self.specs))]

    def at_instances(self):
        return self.split(self.eng.at_nodes().double().cpu().numpy())

    def V_instances(self):
        return self.split(self.eng.V_myo_nodes().double().cpu().numpy())
