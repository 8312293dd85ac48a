"""Small seeded test problems shared by the parity tests."""
from __future__ import annotations

import numpy as np


def slab(nx=20, ny=16, nz=8, pad=1, fibre="rotating", seed=0):
    """A padded slab with a fibre field, phi gradient across z (all three phenotypes)."""
    rng = np.random.default_rng(seed)
    S = np.zeros((nx + 2 * pad, ny + 2 * pad, nz + 2 * pad), dtype=np.uint8)
    S[pad:pad + nx, pad:pad + ny, pad:pad + nz] = 1
    Nx, Ny, Nz = S.shape
    k = np.arange(Nz)[None, None, :] * np.ones((Nx, Ny, 1))
    phi = ((k - pad) / max(1, nz - 1)).astype(np.float32)
    f = np.zeros((Nx, Ny, Nz, 3), dtype=np.float32)
    if fibre == "rotating":
        ang = np.deg2rad(60.0 - 120.0 * np.clip(phi, 0, 1))
        f[..., 0] = np.cos(ang)
        f[..., 1] = np.sin(ang)
        f[..., 2] = 0.15 * rng.standard_normal((Nx, Ny, Nz))
        f /= np.linalg.norm(f, axis=-1, keepdims=True)
    else:
        f[..., 0] = 1.0
    return S, phi, f


def stub_tree(S, voxel_size):
    """Inert 2-node Purkinje stub outside the slab (niederer_common.py:94-111 pattern)."""
    Nx, Ny, Nz = S.shape
    nodes = np.array([[Ny + 5.0, Nx + 5.0, Nz + 5.0], [Ny + 5.0, Nx + 5.0, Nz + 10.0]])
    return nodes, np.array([[0, 1]]), np.array([0.0, 1.0])


def small_tree(S, voxel_size, seed=1, n_branch=6):
    """A short seeded tree hovering over the slab's top face, terminals near the tissue."""
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz = S.shape
    root = np.array([Ny / 2, Nx / 2, Nz + 3.0])
    nodes = [root]
    elements = []
    act = [0.0]
    for b in range(n_branch):
        ang = 2 * np.pi * b / n_branch + 0.2 * rng.standard_normal()
        mid = root + np.array([np.cos(ang) * Ny / 5, np.sin(ang) * Nx / 5, -1.0])
        tip = root + np.array([np.cos(ang) * Ny / 2.6, np.sin(ang) * Nx / 2.6, -(Nz + 2.0)])
        tip[2] = Nz - 1.5
        nodes += [mid, tip]
        elements += [[0, len(nodes) - 2], [len(nodes) - 2, len(nodes) - 1]]
        act += [1.0, 2.0]
    return np.array(nodes), np.array(elements), np.array(act)


def corner_region(S, n=4, pad=1):
    r = np.zeros(S.shape, dtype=bool)
    r[pad:pad + n, pad:pad + n, pad:pad + n] = True
    return r
