"""FP64 CPU oracle for one EMC iteration -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this package, and only as the checker (or as the timed CPU baseline).  The product path
(``emc_amd``) never imports it.

The arithmetic lives in ``emc_oracle.c`` (built to ``liboracle.so`` by ``make -C oracle``); this
module is a thin numpy/ctypes veneer.  Every function names the reference lines it restates in the
C source.  Parity status: every function here is pinned by the checks in ``tests/test_oracle_*.py``
(closed forms, brute force, invariants, independent library routines); the reference's kernels are
CUDA (no nvcc or CUDA headers here), so the reference itself cannot be built as a second checker, and it
ships no tests or golden vectors for this path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_rotation_count.argtypes = [C.c_int]
        L.orc_rotation_count.restype = C.c_int
        L.orc_rotations.argtypes = [C.c_int, f32p, f32p, i32p]
        L.orc_rotations.restype = C.c_int
        L.orc_coords_ewald.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, f32p, f32p, f32p]
        L.orc_coords_ewald.restype = None
        L.orc_coords_flat.argtypes = [C.c_int, f32p, f32p, f32p]
        L.orc_coords_flat.restype = None
        L.orc_mask_annulus.argtypes = [C.c_int, C.c_double, C.c_double, i32p]
        L.orc_mask_annulus.restype = C.c_int
        L.orc_rotmat.argtypes = [f32p, f32p]
        L.orc_rotmat.restype = None
        L.orc_slices.argtypes = [f64p, C.c_int, f32p, C.c_int, f32p, f32p, f32p, C.c_int, C.c_int, f64p]
        L.orc_slices.restype = None
        L.orc_logl.argtypes = [f64p, f64p, i32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                               f64p, C.c_void_p, C.c_double, f64p, C.c_void_p]
        L.orc_logl.restype = None
        L.orc_model_radial.argtypes = [C.c_int, f32p, C.c_int, f64p]
        L.orc_model_radial.restype = None
        L.orc_model_orientations.argtypes = [C.c_int, f32p, C.c_int, i32p, f32p, f32p, f32p, f32p, f64p]
        L.orc_model_orientations.restype = None
        L.orc_normalize_patterns.argtypes = [f32p, C.c_int, C.c_int, i32p, f64p]
        L.orc_normalize_patterns.restype = None
        L.orc_scaling_reference.argtypes = [f32p, f32p, i32p, C.c_int, C.c_int, C.c_int, f64p]
        L.orc_scaling_reference.restype = None
        L.orc_scaling.argtypes = [f64p, f64p, i32p, C.c_int, C.c_int, C.c_int, f64p]
        L.orc_scaling.restype = None
        L.orc_fluence.argtypes = [f64p, f64p, i32p, C.c_int, C.c_int, C.c_int]
        L.orc_fluence.restype = C.c_double
        L.orc_softmax.argtypes = [f64p, C.c_int, C.c_int, f64p]
        L.orc_softmax.restype = C.c_double
        L.orc_mstep.argtypes = [f64p, f64p, i32p, C.c_int, C.c_int, C.c_int, C.c_void_p, f32p, f32p,
                                f32p, f32p, C.c_int, C.c_int, f64p, f64p, C.c_void_p]
        L.orc_mstep.restype = None
        L.orc_compress.argtypes = [f64p, f64p, f64p, C.c_long]
        L.orc_compress.restype = None
        L.orc_blur_taps.argtypes = [C.c_int, C.c_double, f64p]
        L.orc_blur_taps.restype = None
        L.orc_blur.argtypes = [f64p, C.c_int, C.c_double]
        L.orc_blur.restype = None
        L.orc_normalize.argtypes = [f64p, C.c_long, C.POINTER(C.c_double)]
        L.orc_normalize.restype = C.c_long
        L.orc_normalize_support.argtypes = [f64p, f64p, C.c_long, C.POINTER(C.c_double)]
        L.orc_normalize_support.restype = C.c_long
        L.orc_sigma.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int]
        L.orc_sigma.restype = C.c_double
        L.orc_iterate.argtypes = [f64p, C.c_int, f32p, f64p, C.c_int, f64p, i32p, C.c_int, f32p, f32p,
                                  f32p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_iterate.restype = C.c_double
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


DIFF = {"absolute": 0, "poisson": 1, "relative": 2, "true_poisson": 3}
INTERP = {"nearest": 0, "trilinear": 1}


def rotation_count(n: int) -> int:
    return lib().orc_rotation_count(n)


def rotations(n: int):
    """(q [S,4] float32, w [S] float32, stats) restating rotations.c:77-458."""
    S = rotation_count(n)
    q = np.zeros((S, 4), np.float32)
    w = np.zeros(S, np.float32)
    st = np.zeros(5, np.int32)
    r = lib().orc_rotations(n, q, w, st)
    if r != S:
        raise RuntimeError(f"orc_rotations(n={n}) failed: {r} stats={st}")
    return q, w, {"edges": int(st[0]), "faces": int(st[1]), "cells": int(st[2]),
                  "kept": int(st[3]), "total": int(st[4])}


def coords_flat(L: int):
    x, y, z = (np.zeros(L * L, np.float32) for _ in range(3))
    lib().orc_coords_flat(L, x, y, z)
    return x, y, z


def coords_ewald(L: int, pixel_size: float, detector_distance: float, wavelength: float):
    x, y, z = (np.zeros(L * L, np.float32) for _ in range(3))
    lib().orc_coords_ewald(L, pixel_size, detector_distance, wavelength, x, y, z)
    return x, y, z


def mask_annulus(L: int, rb: float, ro: float):
    m = np.zeros(L * L, np.int32)
    lib().orc_mask_annulus(L, rb, ro, m)
    return m


def rotmat(q):
    m = np.zeros(9, np.float32)
    lib().orc_rotmat(np.ascontiguousarray(q, np.float32), m)
    return m.reshape(3, 3)


def slices(model, L, q, coords, interp="trilinear"):
    q = np.ascontiguousarray(q, np.float32).reshape(-1, 4)
    S = q.shape[0]
    x, y, z = (np.ascontiguousarray(c, np.float32) for c in coords)
    P = x.size
    W = np.zeros((S, P), np.float64)
    lib().orc_slices(np.ascontiguousarray(model, np.float64).ravel(), L, q, S, x, y, z, P,
                     INTERP[interp], W)
    return W


def fluence(W, K, mask):
    S, P = W.shape
    I = K.shape[0]
    return lib().orc_fluence(np.ascontiguousarray(W, np.float64), np.ascontiguousarray(K, np.float64),
                             np.ascontiguousarray(mask, np.int32), S, I, P)


def model_radial(L, K):
    """model_input = 3 (emc_helper.cpp:457-500) from patterns K [I][L*L]."""
    K = np.ascontiguousarray(K, np.float32).reshape(-1, L * L)
    out = np.zeros(L ** 3, np.float64)
    lib().orc_model_radial(L, K, K.shape[0], out)
    return out


def model_orientations(L, K, mask, coords, q):
    """model_input = 1 (emc_helper.cpp:501-516) with the orientations q [I][4] given."""
    K = np.ascontiguousarray(K, np.float32).reshape(-1, L * L)
    x, y, z = (np.ascontiguousarray(c, np.float32) for c in coords)
    out = np.zeros(L ** 3, np.float64)
    lib().orc_model_orientations(L, K, K.shape[0], np.ascontiguousarray(mask, np.int32), x, y, z,
                                 np.ascontiguousarray(q, np.float32).reshape(-1, 4), out)
    return out


def normalize_patterns(K, mask):
    """normalize_images (main.cpp:96-97, emc_helper.cpp:389-419)."""
    K = np.ascontiguousarray(K, np.float32)
    out = np.zeros(K.shape, np.float64)
    lib().orc_normalize_patterns(K, K.shape[0], K.shape[1], np.ascontiguousarray(mask, np.int32), out)
    return out


def scaling(W, K, mask):
    """Per (rotation, pattern) intensity scaling phi [S, I] of known_intensity = 0
    (emc_cuda.cu:438-455, 516-531): sum K^2 / sum K W over unmasked pixels with W, K > 1e-10."""
    W = np.ascontiguousarray(W, np.float64)
    K = np.ascontiguousarray(K, np.float64)
    S, P = W.shape
    I = K.shape[0]
    phi = np.zeros((S, I), np.float64)
    lib().orc_scaling(W, K, np.ascontiguousarray(mask, np.int32), S, I, P, phi)
    return phi


def scaling_reference(W, K, mask):
    """The reference's own computation of the same phi (512 threads, int-truncated sums of K W;
    emc_cuda.cu:438-455), for parity studies only."""
    W = np.ascontiguousarray(W, np.float32)
    K = np.ascontiguousarray(K, np.float32)
    S, P = W.shape
    I = K.shape[0]
    phi = np.zeros((S, I), np.float64)
    with np.errstate(all="ignore"):
        lib().orc_scaling_reference(W, K, np.ascontiguousarray(mask, np.int32), S, I, P, phi)
    return phi


def logl(W, K, mask, wrot, diff="true_poisson", sigma=1.0, phi=None, phi0=1.0):
    """Logits R [S, I] (emc_cuda.cu:160-276) and n_s [S]."""
    W = np.ascontiguousarray(W, np.float64)
    K = np.ascontiguousarray(K, np.float64)
    S, P = W.shape
    I = K.shape[0]
    R = np.zeros((S, I), np.float64)
    ns = np.zeros(S, np.int32)
    phi_a = None if phi is None else np.ascontiguousarray(phi, np.float64)
    lib().orc_logl(W, K, np.ascontiguousarray(mask, np.int32), S, I, P, DIFF[diff], sigma,
                   np.ascontiguousarray(wrot, np.float64), _ptr(phi_a), phi0, R, _ptr(ns))
    return R, ns


def softmax(R):
    """(P [S, I], lse [I], total) from logits (emc_cuda.cu:1041-1260, main.cpp:367-392)."""
    Pm = np.array(R, np.float64, order="C", copy=True)
    S, I = Pm.shape
    lse = np.zeros(I, np.float64)
    tot = lib().orc_softmax(Pm, S, I, lse)
    return Pm, lse, tot


def mstep(Presp, K, mask, q, coords, L, interp="trilinear", phi=None):
    """(Mnum, Wt, T) accumulated from zero (emc_atomic.cu:33-132, 183-200)."""
    Presp = np.ascontiguousarray(Presp, np.float64)
    K = np.ascontiguousarray(K, np.float64)
    S, I = Presp.shape
    P = K.shape[1]
    q = np.ascontiguousarray(q, np.float32).reshape(-1, 4)
    x, y, z = (np.ascontiguousarray(c, np.float32) for c in coords)
    Mnum = np.zeros(L ** 3, np.float64)
    Wt = np.zeros(L ** 3, np.float64)
    T = np.zeros(S, np.float64)
    phi_a = None if phi is None else np.ascontiguousarray(phi, np.float64)
    lib().orc_mstep(Presp, K, np.ascontiguousarray(mask, np.int32), S, I, P, _ptr(phi_a), q, x, y, z,
                    L, INTERP[interp], Mnum, Wt, _ptr(T))
    return Mnum, Wt, T


def compress(Mnum, Wt):
    M = np.zeros_like(Mnum, dtype=np.float64)
    lib().orc_compress(M, np.ascontiguousarray(Mnum, np.float64), np.ascontiguousarray(Wt, np.float64),
                       M.size)
    return M


def blur_taps(L, sigma):
    g = np.zeros(L, np.float64)
    lib().orc_blur_taps(L, sigma, g)
    return g


def blur(M, L, sigma):
    out = np.array(M, np.float64, copy=True).ravel()
    lib().orc_blur(out, L, sigma)
    return out


def normalize(M):
    out = np.array(M, np.float64, copy=True).ravel()
    avg = C.c_double(0.0)
    cnt = lib().orc_normalize(out, out.size, C.byref(avg))
    return out, avg.value, cnt


def normalize_support(M, Wt):
    """The iteration's normaliser: M / mean(M over Wt > 0), on every voxel."""
    out = np.array(M, np.float64, copy=True).ravel()
    avg = C.c_double(0.0)
    cnt = lib().orc_normalize_support(out, np.ascontiguousarray(Wt, np.float64).ravel(), out.size, C.byref(avg))
    return out, avg.value, cnt


def sigma_at(iteration, s_start, s_final, half_life):
    return lib().orc_sigma(iteration, s_start, s_final, half_life)


def iterate(model, L, q, w, K, mask, coords, interp="trilinear", diff="true_poisson", sigma=1.0,
            fluence_value=0.0, blur_sigma=0.0, want_P=False, scaling=False):
    """One EMC iteration (main.cpp:239-480 order).  Returns dict with model, Wt, lse, total, ..."""
    model = np.array(model, np.float64, copy=True).ravel()
    q = np.ascontiguousarray(q, np.float32).reshape(-1, 4)
    S = q.shape[0]
    K = np.ascontiguousarray(K, np.float64)
    I = K.shape[0]
    x, y, z = (np.ascontiguousarray(c, np.float32) for c in coords)
    Wt = np.zeros(L ** 3, np.float64)
    lse = np.zeros(I, np.float64)
    Pm = np.zeros((S, I), np.float64) if want_P else None
    stats = np.zeros(2, np.float64)
    tot = lib().orc_iterate(model, L, q, np.ascontiguousarray(w, np.float64), S, K,
                            np.ascontiguousarray(mask, np.int32), I, x, y, z, INTERP[interp],
                            DIFF[diff], sigma, fluence_value, blur_sigma, int(bool(scaling)),
                            _ptr(Wt), _ptr(Pm),
                            _ptr(lse), _ptr(stats))
    return {"model": model, "Wt": Wt, "lse": lse, "total": tot, "P": Pm, "phi0": stats[0],
            "avg": stats[1]}


# ---- orientation-proposal MLP (emc_oracle_mlp.c) ----------------------------------------------

def _mlp_sigs(L):
    L.orc_mlp_forward.argtypes = [f64p, C.c_int, C.c_int, C.c_int, C.c_int] + [f64p] * 6 + [f64p] * 6
    L.orc_mlp_forward.restype = None
    L.orc_mlp_loss.argtypes = [f64p, f64p, C.c_int, C.c_int]
    L.orc_mlp_loss.restype = C.c_double
    L.orc_mlp_backward.argtypes = [f64p, f64p, C.c_int, C.c_int, C.c_int, C.c_int] + [f64p] * 2 + [f64p] * 5 + \
        [f64p] * 6
    L.orc_mlp_backward.restype = None
    L.orc_adam.argtypes = [f64p, f64p, f64p, f64p, C.c_long, C.c_int, C.c_double, C.c_double, C.c_double,
                           C.c_double]
    L.orc_adam.restype = None


def mlp_forward(x, params):
    """params: dict W1 [H,D], b1, W2 [H,H], b2, W3 [C,H], b3.  Returns dict of activations."""
    L = lib()
    if not hasattr(L, "_mlp"):
        _mlp_sigs(L)
        L._mlp = True
    x = np.ascontiguousarray(x, np.float64)
    B, D = x.shape
    H = params["W1"].shape[0]
    Cc = params["W3"].shape[0]
    a = {k: np.zeros((B, H)) for k in ("pre1", "h1", "pre2", "h2")}
    a["z"] = np.zeros((B, Cc))
    a["p"] = np.zeros((B, Cc))
    P = {k: np.ascontiguousarray(v, np.float64) for k, v in params.items()}
    L.orc_mlp_forward(x, B, D, H, Cc, P["W1"], P["b1"], P["W2"], P["b2"], P["W3"], P["b3"], a["pre1"], a["h1"],
                      a["pre2"], a["h2"], a["z"], a["p"])
    return a


def mlp_loss(p, t):
    L = lib()
    if not hasattr(L, "_mlp"):
        _mlp_sigs(L)
        L._mlp = True
    p = np.ascontiguousarray(p, np.float64)
    return L.orc_mlp_loss(p, np.ascontiguousarray(t, np.float64), p.shape[0], p.shape[1])


def mlp_backward(x, t, params, acts):
    L = lib()
    if not hasattr(L, "_mlp"):
        _mlp_sigs(L)
        L._mlp = True
    x = np.ascontiguousarray(x, np.float64)
    B, D = x.shape
    H = params["W1"].shape[0]
    Cc = params["W3"].shape[0]
    g = {k: np.zeros_like(np.asarray(v, np.float64)) for k, v in params.items()}
    P = {k: np.ascontiguousarray(v, np.float64) for k, v in params.items()}
    L.orc_mlp_backward(x, np.ascontiguousarray(t, np.float64), B, D, H, Cc, P["W2"], P["W3"], acts["pre1"],
                       acts["h1"], acts["pre2"], acts["h2"], acts["p"], g["W1"], g["b1"], g["W2"], g["b2"],
                       g["W3"], g["b3"])
    return g


def adam(w, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """In-place Adam step t (1-based) on float64 arrays."""
    L = lib()
    if not hasattr(L, "_mlp"):
        _mlp_sigs(L)
        L._mlp = True
    L.orc_adam(w, np.ascontiguousarray(g, np.float64), m, v, w.size, t, lr, b1, b2, eps)
