"""ctypes binding of ``libemc.so`` (declared in ``include/emc.h``).

There is no fallback: if the HIP library is not built, importing the compute entry points raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libemc.so")

EMC_OK = 0
STATUS_NAMES = {0: "EMC_OK", 1: "EMC_E_INVALID", 2: "EMC_E_HIP", 3: "EMC_E_OOM", 4: "EMC_E_NONFINITE",
                5: "EMC_E_UNSUPPORTED", 6: "EMC_E_STATE", 7: "EMC_E_IO"}

DIFF = {"absolute": 0, "poisson": 1, "relative": 2, "true_poisson": 3}
INTERP = {"nearest": 0, "trilinear": 1}
PRECISION = {"fp32": 0, "bf16x2": 1, "bf16x3": 2}
INSERT = {"atomic": 0, "fused": 1, "binned": 2}
MSTEP = {"auto": 0, "dense": 1, "sparse": 2}
N_STAGES = 10
STAGES = ["slice", "logl", "softmax", "mstep", "compress", "total", "logl_gemm", "mstep_gemm", "propose", "train"]
SEARCH = {0: "full", 1: "proposal"}


class Dims(C.Structure):
    _fields_ = [("model_side", C.c_int32), ("n_rotations", C.c_int32), ("n_patterns", C.c_int32),
                ("slice_chunk", C.c_int32)]


class Options(C.Structure):
    _fields_ = [("diff", C.c_int32), ("interp", C.c_int32), ("precision", C.c_int32), ("insert", C.c_int32),
                ("sigma_start", C.c_double), ("sigma_final", C.c_double), ("sigma_half_life", C.c_int32),
                ("device", C.c_int32), ("blur_sigma", C.c_double), ("fluence", C.c_double),
                ("timing", C.c_int32), ("mfma_tile", C.c_int32), ("count_bytes", C.c_int32),
                ("mstep", C.c_int32), ("scaling", C.c_int32), ("tile_rows", C.c_int32),
                ("mfma_staging", C.c_int32), ("keep_logits", C.c_int32)]


class IterStats(C.Structure):
    _fields_ = [("total_logl", C.c_double), ("sigma", C.c_double), ("fluence", C.c_double),
                ("model_avg", C.c_double), ("weight_sum", C.c_double), ("stage_ms", C.c_float * N_STAGES),
                ("search", C.c_int32), ("n_fallback", C.c_int32), ("predicted_mass", C.c_double),
                ("captured_mass", C.c_double), ("train_loss", C.c_double), ("work_fraction", C.c_double),
                ("logits_skipped", C.c_double)]


class ProposalOptions(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("top_k", "warmup_full", "full_every", "hidden", "epochs", "batch",
                                          "synth_patterns", "refresh_epochs", "recall", "train_patterns")] + \
               [(n, C.c_float) for n in ("min_mass", "lr", "max_fallback", "refresh_fraction", "train_below")] + \
               [("seed", C.c_uint64)]


class Config(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("model_side", "read_stride", "detector_size", "rotations_n",
                                          "sigma_half_life", "slice_chunk", "N_images", "max_iterations")] + \
               [(n, C.c_double) for n in ("wavelength", "pixel_size", "detector_distance", "sigma_start",
                                           "sigma_final")] + \
               [(n, C.c_int32) for n in ("diff", "normalize_images", "known_intensity", "model_input",
                                          "exclude_images", "output_period", "output_loop", "isOutput", "isDebug",
                                          "blur_image", "nGPU")] + \
               [(n, C.c_double) for n in ("blur_sigma", "exclude_ratio", "model_blur")] + \
               [(n, C.c_char * 512) for n in ("image_prefix", "mask_file", "model_file")] + \
               [(n, C.c_int32) for n in ("precision", "interp", "insert", "flat_detector")] + \
               [("fluence", C.c_double), ("synth_atoms", C.c_int32)] + \
               [(n, C.c_double) for n in ("synth_radius", "synth_photons", "beamstop", "outer_radius")] + \
               [("synth_seed", C.c_uint64), ("proposal_coarse_n", C.c_int32), ("proposal_top_k", C.c_int32)]


class EmcError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


_lib = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
vp = C.c_void_p
ctxp = C.c_void_p

# name -> (restype, argtypes); every symbol of include/emc.h
SIGNATURES = {
    "emc_default_options": (None, [C.POINTER(Options)]),
    "emc_create": (C.c_int, [C.POINTER(Dims), C.POINTER(Options), C.POINTER(ctxp)]),
    "emc_destroy": (None, [ctxp]),
    "emc_last_error": (C.c_char_p, [ctxp]),
    "emc_device_info": (C.c_int, [C.c_int32, C.c_char_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "emc_set_rotations": (C.c_int, [ctxp, f32p, f32p]),
    "emc_set_coords": (C.c_int, [ctxp, f32p, f32p, f32p]),
    "emc_set_mask": (C.c_int, [ctxp, i32p]),
    "emc_set_patterns_i32": (C.c_int, [ctxp, i32p]),
    "emc_set_patterns_f32": (C.c_int, [ctxp, f32p]),
    "emc_set_model": (C.c_int, [ctxp, f32p]),
    "emc_get_model": (C.c_int, [ctxp, vp, vp]),
    "emc_model_stats": (C.c_int, [ctxp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "emc_get_logl": (C.c_int, [ctxp, f32p]),
    "emc_get_resp": (C.c_int, [ctxp, f32p]),
    "emc_get_logl_block": (C.c_int, [ctxp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, f32p]),
    "emc_get_resp_block": (C.c_int, [ctxp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, f32p]),
    "emc_get_lse": (C.c_int, [ctxp, f64p]),
    "emc_get_slice_weights": (C.c_int, [ctxp, f64p]),
    "emc_expand": (C.c_int, [ctxp, C.c_int32, C.c_int32, f32p]),
    "emc_logl": (C.c_int, [ctxp, C.c_float]),
    "emc_normalize_resp": (C.c_int, [ctxp, C.POINTER(C.c_double)]),
    "emc_set_resp": (C.c_int, [ctxp, f32p]),
    "emc_set_mstep_sums": (C.c_int, [ctxp, f32p, f32p]),
    "emc_maximize": (C.c_int, [ctxp]),
    "emc_compress": (C.c_int, [ctxp, C.POINTER(C.c_double)]),
    "emc_normalize_model": (C.c_int, [ctxp, C.POINTER(C.c_double)]),
    "emc_iterate": (C.c_int, [ctxp, C.c_int32, C.POINTER(IterStats)]),
    "emc_compress_device": (C.c_int, [C.c_int32, C.c_double, vp, vp, vp, vp, vp]),
    "emc_rotation_count": (C.c_int32, [C.c_int32]),
    "emc_generate_rotations": (C.c_int, [C.c_int32, f32p, f32p]),
    "emc_coords_ewald": (C.c_int, [C.c_int32, C.c_float, C.c_float, C.c_float, f32p, f32p, f32p]),
    "emc_coords_flat": (C.c_int, [C.c_int32, f32p, f32p, f32p]),
    "emc_mask_annulus": (C.c_int32, [C.c_int32, C.c_double, C.c_double, i32p]),
    "emc_sigma_schedule": (C.c_double, [C.c_int32, C.c_double, C.c_double, C.c_int32]),
    "emc_synth_object": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_uint64, f32p]),
    "emc_synth_model": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_uint64, f32p]),
    "emc_synth_patterns": (C.c_int, [f32p, C.c_int32, f32p, f32p, f32p, i32p, C.c_int32, C.c_double, C.c_uint64,
                                     i32p, vp]),
    "emc_synth_uniform_model": (C.c_int, [C.c_int32, C.c_uint64, f32p]),
    "emc_model_init_radial": (C.c_int, [C.c_int32, f32p, C.c_int32, f32p]),
    "emc_model_init_orientations": (C.c_int, [C.c_int32, f32p, C.c_int32, i32p, f32p, f32p, f32p, C.c_uint64,
                                              C.c_void_p, f32p]),
    "emc_model_read_raw": (C.c_int, [C.c_char_p, C.c_int32, f32p]),
    "emc_normalize_patterns": (C.c_int, [f32p, C.c_int32, C.c_int32, i32p, f32p]),
    "emc_config_defaults": (C.c_int, [C.POINTER(Config)]),
    "emc_config_parse": (C.c_int, [C.c_char_p, C.POINTER(Config)]),
    "emc_config_read": (C.c_int, [C.c_char_p, C.POINTER(Config)]),
    "emc_proposal_defaults": (None, [C.POINTER(ProposalOptions)]),
    "emc_coarse_partition": (C.c_int, [f32p, C.c_int32, f32p, C.c_int32, i32p]),
    "emc_proposal_enable": (C.c_int, [ctxp, f32p, C.c_int32, C.POINTER(ProposalOptions)]),
    "emc_proposal_disable": (C.c_int, [ctxp]),
    "emc_proposal_predict": (C.c_int, [ctxp, vp, vp]),
    "emc_proposal_recall": (C.c_int, [ctxp, vp]),
    "emc_proposal_cell_targets": (C.c_int, [ctxp, f32p]),
    "emc_proposal_captured": (C.c_int, [ctxp, f64p]),
    "emc_proposal_train": (C.c_int, [ctxp, C.c_int32, C.POINTER(C.c_double)]),
    "emc_proposal_features": (C.c_int, [ctxp, f32p]),
    "emc_mlp_get_params": (C.c_int, [ctxp, vp, vp, vp, vp, vp, vp]),
    "emc_mlp_set_params": (C.c_int, [ctxp, f32p, f32p, f32p, f32p, f32p, f32p]),
    "emc_mlp_forward": (C.c_int, [ctxp, f32p, C.c_int32, f32p]),
    "emc_mlp_gradients": (C.c_int, [ctxp, f32p, f32p, C.c_int32, C.POINTER(C.c_double), vp, vp, vp, vp, vp, vp]),
    "emc_mlp_train_step": (C.c_int, [ctxp, f32p, f32p, C.c_int32, C.POINTER(C.c_double)]),
}


def lib():
    """Load libemc.so (built in-tree by ``make -C emc_amd/csrc``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                              "or `make -C emc_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, ctx=None):
    if status != EMC_OK:
        msg = lib().emc_last_error(ctx)
        raise EmcError(status, msg.decode() if msg else "")
