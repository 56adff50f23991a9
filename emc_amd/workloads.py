"""Synthetic EMC workloads (fixed seeds), named by their sizes.

Each workload: Gaussian atoms (sigma 1 voxel) uniform in a sphere -> model = fftshifted |DFT|^2,
mean 1; patterns = trilinear flat-plane slices at seeded uniform random rotations, scaled to the
mean photon count over the unmasked annulus, Poisson-sampled to int32 counts; initial model
uniform [0, 1) (then normalised as main.cpp:180 does); no per-pattern scaling.
"""
from __future__ import annotations

import numpy as np

from . import coords_flat, mask_annulus, rotations, synth_model, synth_patterns, synth_uniform_model

WORKLOADS = {
    # 32^3 model, 3,240 rotations (n=4), 1,000 patterns of 32x32, 2,000 photons; 10 iterations
    "L32_S3240_I1000": dict(L=32, n_atoms=40, radius=6.0, rot_n=4, I=1000, photons=2000.0, rb=3.0, ro=16.0,
                            iters=10, blur=0.5, seed=11),
    # 64^3 model, 25,680 rotations (n=8), 20,000 patterns of 64x64, 5,000 photons; 5 iterations
    "L64_S25680_I20000": dict(L=64, n_atoms=200, radius=12.0, rot_n=8, I=20000, photons=5000.0, rb=4.0,
                              ro=32.0, iters=5, blur=0.5, seed=12),
    # 128^3 model, 50,100 rotations (n=10), 50,000 patterns of 128x128, 300 photons (sparse counts)
    "L128_S50100_I50000": dict(L=128, n_atoms=800, radius=24.0, rot_n=10, I=50000, photons=300.0, rb=4.0,
                               ro=64.0, iters=3, blur=0.5, seed=13),
    # 64^3 model (the 64^3 object), 204,960 fine rotations (n=16), 50,000 patterns, 2,000 photons
    "L64_S204960_I50000": dict(L=64, n_atoms=200, radius=12.0, rot_n=16, I=50000, photons=2000.0, rb=4.0,
                               ro=32.0, iters=5, blur=0.5, seed=12, coarse_n=2),
}


def make(name: str, I: int | None = None, rot_n: int | None = None, with_rotations: bool = True,
         photons: float | None = None) -> dict:
    """Arrays of a workload; I / rot_n / photons override the pattern count / rotation sampling /
    mean photons per pattern."""
    cfg = dict(WORKLOADS[name])
    if I is not None:
        cfg["I"] = I
    if rot_n is not None:
        cfg["rot_n"] = rot_n
    if photons is not None:
        cfg["photons"] = photons
    L = cfg["L"]
    seed = cfg["seed"]
    model_true = synth_model(L, cfg["n_atoms"], cfg["radius"], 1000 + seed)
    coords = coords_flat(L)
    mask = mask_annulus(L, cfg["rb"], cfg["ro"])
    K, q_true = synth_patterns(model_true, L, coords, mask, cfg["I"], cfg["photons"], 2000 + seed)
    model0 = synth_uniform_model(L, 3000 + seed)
    out = dict(cfg=cfg, L=L, coords=coords, mask=mask, K=K, q_true=q_true, model_true=model_true, model0=model0)
    if with_rotations:
        q, w = rotations(cfg["rot_n"])
        out.update(q=q, w=w)
    return out


def normalized(model: np.ndarray) -> np.ndarray:
    """model / mean(model > 0) on every voxel (emc_cuda.cu:788-798), in FP64."""
    m = np.asarray(model, np.float64)
    pos = m[m > 0]
    return m / pos.mean()
