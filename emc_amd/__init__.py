"""emc_amd -- MI355X-native EMC (Expand-Maximize-Compress) reconstruction engine.

The compute path is ``libemc.so``: hand-written gfx950 HIP kernels behind the C ABI of
``include/emc.h``.  This package is a thin Python veneer over that ABI (ctypes) for tests, the
benchmark and scripting; it mirrors the reference's stage API (emc_cuda_help.h:9-117, driven by
main.cpp:239-516): slices -> likelihoods -> responsibilities -> maximize -> compress.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (DIFF, INSERT, INTERP, MSTEP, PRECISION, SEARCH, STAGES, Dims, EmcError, IterStats, Options,
                   ProposalOptions, check, lib)

__all__ = ["Engine", "EmcError", "rotations", "rotation_count", "coords_flat", "coords_ewald", "mask_annulus",
           "model_init_radial", "model_init_orientations", "model_read_raw", "normalize_patterns",
           "sigma_schedule", "synth_object", "synth_model", "synth_patterns", "synth_uniform_model",
           "default_options", "device_info", "proposal_options", "coarse_partition"]


# ---- host utilities (no GPU) -------------------------------------------------------------------

def rotation_count(n: int) -> int:
    return int(lib().emc_rotation_count(n))


def rotations(n: int):
    """Refined 600-cell rotation list (q [S,4], w [S]) -- generate_rotation_list (rotations.c:77)."""
    S = rotation_count(n)
    if S <= 0:
        raise ValueError(f"bad rotations_n {n}")
    q = np.zeros((S, 4), np.float32)
    w = np.zeros(S, np.float32)
    check(lib().emc_generate_rotations(n, q, w))
    return q, w


def coords_flat(L: int):
    x, y, z = (np.zeros(L * L, np.float32) for _ in range(3))
    check(lib().emc_coords_flat(L, x, y, z))
    return x, y, z


def coords_ewald(L: int, pixel_size: float, detector_distance: float, wavelength: float):
    """emc_helper::calculate_coordinates (emc_helper.cpp:55-79)."""
    x, y, z = (np.zeros(L * L, np.float32) for _ in range(3))
    check(lib().emc_coords_ewald(L, pixel_size, detector_distance, wavelength, x, y, z))
    return x, y, z


def mask_annulus(L: int, rb: float, ro: float):
    m = np.zeros(L * L, np.int32)
    n = lib().emc_mask_annulus(L, rb, ro, m)
    if n < 0:
        raise ValueError("mask_annulus failed")
    return m


def sigma_schedule(iteration: int, s_start: float, s_final: float, half_life: int) -> float:
    return float(lib().emc_sigma_schedule(iteration, s_start, s_final, half_life))


def synth_object(L: int, n_atoms: int, radius: float, seed: int):
    o = np.zeros(L ** 3, np.float32)
    check(lib().emc_synth_object(L, n_atoms, radius, seed, o))
    return o


def synth_model(L: int, n_atoms: int, radius: float, seed: int):
    m = np.zeros(L ** 3, np.float32)
    check(lib().emc_synth_model(L, n_atoms, radius, seed, m))
    return m


def synth_patterns(model, L: int, coords, mask, n_patterns: int, photons: float, seed: int):
    x, y, z = (np.ascontiguousarray(c, np.float32) for c in coords)
    K = np.zeros((n_patterns, L * L), np.int32)
    q = np.zeros((n_patterns, 4), np.float32)
    check(lib().emc_synth_patterns(np.ascontiguousarray(model, np.float32), L, x, y, z,
                                   np.ascontiguousarray(mask, np.int32), n_patterns, photons, seed, K,
                                   q.ctypes.data_as(C.c_void_p)))
    return K, q


def synth_uniform_model(L: int, seed: int):
    m = np.zeros(L ** 3, np.float32)
    check(lib().emc_synth_uniform_model(L, seed, m))
    return m


def default_options(**kw) -> Options:
    o = Options()
    lib().emc_default_options(C.byref(o))
    for k, v in kw.items():
        if k == "diff":
            v = DIFF[v] if isinstance(v, str) else v
        elif k == "interp":
            v = INTERP[v] if isinstance(v, str) else v
        elif k == "precision":
            v = PRECISION[v] if isinstance(v, str) else v
        elif k == "insert":
            v = INSERT[v] if isinstance(v, str) else v
        elif k == "mstep":
            v = MSTEP[v] if isinstance(v, str) else v
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def model_init_radial(L: int, patterns):
    """model_input = 3: radial average of the patterns [I][L*L] (emc_model_init_radial)."""
    K = np.ascontiguousarray(patterns, np.float32).reshape(-1, L * L)
    m = np.zeros(L ** 3, np.float32)
    check(lib().emc_model_init_radial(L, K, K.shape[0], m))
    return m


def model_init_orientations(L: int, patterns, mask, coords, seed: int):
    """model_input = 1: patterns inserted at seeded random orientations; returns (model, q [I][4])."""
    K = np.ascontiguousarray(patterns, np.float32).reshape(-1, L * L)
    x, y, z = (np.ascontiguousarray(c, np.float32) for c in coords)
    q = np.zeros((K.shape[0], 4), np.float32)
    m = np.zeros(L ** 3, np.float32)
    check(lib().emc_model_init_orientations(L, K, K.shape[0], np.ascontiguousarray(mask, np.int32), x, y, z, seed,
                                            q.ctypes.data_as(C.c_void_p), m))
    return m, q


def model_read_raw(path: str, L: int):
    """model_input = 2: |model| from raw float32 [L^3] (emc_model_read_raw)."""
    m = np.zeros(L ** 3, np.float32)
    check(lib().emc_model_read_raw(str(path).encode(), L, m))
    return m


def normalize_patterns(patterns, mask):
    """normalize_images: each pattern scaled by P / its sum over mask == 1 (emc_normalize_patterns)."""
    K = np.ascontiguousarray(patterns, np.float32)
    out = np.zeros_like(K)
    check(lib().emc_normalize_patterns(K, K.shape[0], K.shape[1], np.ascontiguousarray(mask, np.int32), out))
    return out


def device_info(device: int = 0):
    name = C.create_string_buffer(256)
    cu = C.c_int32(0)
    mem = C.c_int64(0)
    check(lib().emc_device_info(device, name, 256, C.byref(cu), C.byref(mem)))
    return {"name": name.value.decode(), "n_cu": cu.value, "hbm_bytes": mem.value}


# ---- engine ----------------------------------------------------------------------------------

class Engine:
    """One device context (include/emc.h ``emc_ctx``).  All arrays are host numpy arrays."""

    def __init__(self, L: int, S: int, I: int, slice_chunk: int = 0, **options):
        self.L, self.S, self.I = L, S, I
        self.n_cells = self.top_k = self.hidden = self.recall = 0  # set by enable_proposal
        self.opt = default_options(**options)
        d = Dims(L, S, I, slice_chunk)
        h = C.c_void_p()
        check(lib().emc_create(C.byref(d), C.byref(self.opt), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().emc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, st):
        check(st, self._h)

    # inputs
    def set_rotations(self, q, w):
        self._c(lib().emc_set_rotations(self._h, np.ascontiguousarray(q, np.float32).reshape(-1),
                                        np.ascontiguousarray(w, np.float32)))

    def set_coords(self, coords):
        x, y, z = (np.ascontiguousarray(c, np.float32) for c in coords)
        self._c(lib().emc_set_coords(self._h, x, y, z))

    def set_mask(self, mask):
        self._c(lib().emc_set_mask(self._h, np.ascontiguousarray(mask, np.int32)))

    def set_patterns(self, K):
        K = np.asarray(K)
        if K.dtype == np.int32:
            self._c(lib().emc_set_patterns_i32(self._h, np.ascontiguousarray(K).reshape(-1)))
        else:
            self._c(lib().emc_set_patterns_f32(self._h, np.ascontiguousarray(K, np.float32).reshape(-1)))

    def set_model(self, model):
        self._c(lib().emc_set_model(self._h, np.ascontiguousarray(model, np.float32).reshape(-1)))

    def setup(self, q, w, coords, mask, K, model):
        self.set_rotations(q, w)
        self.set_coords(coords)
        self.set_mask(mask)
        self.set_patterns(K)
        self.set_model(model)

    # outputs
    def get_model(self):
        M = np.zeros(self.L ** 3, np.float32)
        Wt = np.zeros(self.L ** 3, np.float32)
        self._c(lib().emc_get_model(self._h, M.ctypes.data_as(C.c_void_p), Wt.ctypes.data_as(C.c_void_p)))
        return M, Wt

    def model_stats(self):
        mx, avg = C.c_double(), C.c_double()
        self._c(lib().emc_model_stats(self._h, C.byref(mx), C.byref(avg)))
        return mx.value, avg.value

    def expand(self, s0: int, count: int):
        out = np.zeros((count, self.L * self.L), np.float32)
        self._c(lib().emc_expand(self._h, s0, count, out))
        return out

    def get_logl(self):
        R = np.zeros((self.S, self.I), np.float32)
        self._c(lib().emc_get_logl(self._h, R))
        return R

    def get_resp(self):
        P = np.zeros((self.S, self.I), np.float32)
        self._c(lib().emc_get_resp(self._h, P))
        return P

    def get_logl_block(self, s0: int, ns: int, i0: int, ni: int):
        """Logits of rotations [s0, s0+ns) and patterns [i0, i0+ni): [ns][ni]."""
        R = np.zeros((ns, ni), np.float32)
        self._c(lib().emc_get_logl_block(self._h, s0, ns, i0, ni, R))
        return R

    def get_resp_block(self, s0: int, ns: int, i0: int, ni: int):
        P = np.zeros((ns, ni), np.float32)
        self._c(lib().emc_get_resp_block(self._h, s0, ns, i0, ni, P))
        return P

    def get_lse(self):
        l = np.zeros(self.I, np.float64)
        self._c(lib().emc_get_lse(self._h, l))
        return l

    def get_slice_weights(self):
        T = np.zeros(self.S, np.float64)
        self._c(lib().emc_get_slice_weights(self._h, T))
        return T

    # stages
    def logl(self, sigma: float = 1.0):
        self._c(lib().emc_logl(self._h, sigma))

    def normalize_resp(self) -> float:
        t = C.c_double()
        self._c(lib().emc_normalize_resp(self._h, C.byref(t)))
        return t.value

    def set_resp(self, P):
        self._c(lib().emc_set_resp(self._h, np.ascontiguousarray(P, np.float32).reshape(-1)))

    def set_mstep_sums(self, numerator, weight):
        """Replace the M-step sums (model numerator and weight, [L^3]) that compress divides."""
        self._c(lib().emc_set_mstep_sums(self._h, np.ascontiguousarray(numerator, np.float32).reshape(-1),
                                         np.ascontiguousarray(weight, np.float32).reshape(-1)))

    def maximize(self):
        self._c(lib().emc_maximize(self._h))

    def compress(self) -> float:
        a = C.c_double()
        self._c(lib().emc_compress(self._h, C.byref(a)))
        return a.value

    def normalize_model(self) -> float:
        a = C.c_double()
        self._c(lib().emc_normalize_model(self._h, C.byref(a)))
        return a.value

    def iterate(self, iteration: int = 0) -> dict:
        st = IterStats()
        self._c(lib().emc_iterate(self._h, iteration, C.byref(st)))
        return {"total_logl": st.total_logl, "sigma": st.sigma, "fluence": st.fluence, "model_avg": st.model_avg,
                "weight_sum": st.weight_sum, "stage_ms": {k: float(st.stage_ms[i]) for i, k in enumerate(STAGES)},
                "search": SEARCH.get(st.search, st.search), "n_fallback": st.n_fallback,
                "predicted_mass": st.predicted_mass, "captured_mass": st.captured_mass,
                "train_loss": st.train_loss, "work_fraction": st.work_fraction,
                "logits_skipped": st.logits_skipped}

    # ---- learned orientation proposal (no reference counterpart) ----

    def enable_proposal(self, coarse_q, **options):
        """Partition this engine's rotations into the cells of ``coarse_q`` [C][4] and set up the
        network; options are fields of ``emc_proposal_options``."""
        o = proposal_options(**options)
        q = np.ascontiguousarray(coarse_q, np.float32)
        self.n_cells = q.shape[0]
        self.top_k = o.top_k
        self.hidden = o.hidden
        self.recall = o.recall
        self._c(lib().emc_proposal_enable(self._h, q.reshape(-1), q.shape[0], C.byref(o)))

    def disable_proposal(self):
        self._c(lib().emc_proposal_disable(self._h))

    def proposal_predict(self):
        cells = np.zeros((self.I, self.top_k), np.int32)
        mass = np.zeros(self.I, np.float32)
        self._c(lib().emc_proposal_predict(self._h, cells.ctypes.data, mass.ctypes.data))
        return cells, mass

    def proposal_captured(self):
        """Per-pattern captured mass of the last full-search iteration that measured it."""
        cap = np.zeros(self.I, np.float64)
        self._c(lib().emc_proposal_captured(self._h, cap))
        return cap

    def proposal_recall(self):
        """Recalled cells [I][recall] (most massive first, -1 = none)."""
        cells = np.full((self.I, self.recall), -1, np.int32)
        self._c(lib().emc_proposal_recall(self._h, cells.ctypes.data))
        return cells

    def proposal_cell_targets(self):
        t = np.zeros((self.I, self.n_cells), np.float32)
        self._c(lib().emc_proposal_cell_targets(self._h, t))
        return t

    def proposal_train(self, epochs: int) -> float:
        loss = C.c_double()
        self._c(lib().emc_proposal_train(self._h, epochs, C.byref(loss)))
        return loss.value

    def proposal_features(self):
        raw = np.zeros((self.I, 1024), np.float32)
        self._c(lib().emc_proposal_features(self._h, raw))
        return raw

    def mlp_params(self):
        H, Cn = self.hidden, self.n_cells
        out = [np.zeros(s, np.float32) for s in ((H, 1024), H, (H, H), H, (Cn, H), Cn)]
        self._c(lib().emc_mlp_get_params(self._h, *[a.ctypes.data for a in out]))
        return out

    def set_mlp_params(self, W1, b1, W2, b2, W3, b3):
        arrs = [np.ascontiguousarray(a, np.float32).reshape(-1) for a in (W1, b1, W2, b2, W3, b3)]
        self._c(lib().emc_mlp_set_params(self._h, *arrs))

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        probs = np.zeros((x.shape[0], self.n_cells), np.float32)
        self._c(lib().emc_mlp_forward(self._h, x.reshape(-1), x.shape[0], probs))
        return probs

    def mlp_gradients(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        H, Cn = self.hidden, self.n_cells
        g = [np.zeros(s, np.float32) for s in ((H, 1024), H, (H, H), H, (Cn, H), Cn)]
        loss = C.c_double()
        self._c(lib().emc_mlp_gradients(self._h, x.reshape(-1), t.reshape(-1), x.shape[0], C.byref(loss),
                                        *[a.ctypes.data for a in g]))
        return loss.value, g

    def mlp_train_step(self, x, t) -> float:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_double()
        self._c(lib().emc_mlp_train_step(self._h, x.reshape(-1), t.reshape(-1), x.shape[0], C.byref(loss)))
        return loss.value


def proposal_options(**kw) -> ProposalOptions:
    o = ProposalOptions()
    lib().emc_proposal_defaults(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k) or k.startswith("reserved"):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def coarse_partition(fine_q, coarse_q):
    """Cell of each fine rotation: argmax_c |q_s . q_c| (lowest c on ties)."""
    f = np.ascontiguousarray(fine_q, np.float32)
    q = np.ascontiguousarray(coarse_q, np.float32)
    cell = np.zeros(f.shape[0], np.int32)
    check(lib().emc_coarse_partition(f.reshape(-1), f.shape[0], q.reshape(-1), q.shape[0], cell))
    return cell
