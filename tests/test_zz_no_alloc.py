"""The allocation-free entry point emc_compress_device (static per-device scratch) against the
existing emc_compress on the same M-step sums, through the C ABI: two calls in a row, the second
with a smaller model; and its rejection of a model too large for the scratch."""
import ctypes as C

import numpy as np
import pytest

import emc_amd
from emc_amd._lib import EmcError, check, lib

pytestmark = pytest.mark.gpu


def sums(L, seed):
    rng = np.random.default_rng(seed)
    Wt = rng.uniform(0.0, 3.0, L ** 3).astype(np.float32)
    Wt[rng.random(L ** 3) < 0.2] = 0.0                      # voxels without weight: missing
    Mnum = (rng.uniform(0.0, 2.0, L ** 3) * Wt).astype(np.float32)
    Mnum[rng.random(L ** 3) < 0.1] = 0.0                    # weight but no counts
    return Mnum, Wt


def existing(L, Mnum, Wt, sigma):
    with emc_amd.Engine(L, 4, 4, blur_sigma=sigma) as e:
        e.set_mstep_sums(Mnum, Wt)
        avg = e.compress()
        return e.get_model()[0], avg


def hip_runtime():
    """The HIP runtime libemc.so is linked against (already mapped into this process)."""
    lib()
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64" in ln}
    assert paths, "libamdhip64 is not loaded"
    h = C.CDLL(sorted(paths)[0])
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def test_compress_device_equals_compress_twice():
    hip = hip_runtime()
    bufs = []

    def dev_alloc(nbytes):
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), nbytes) == 0
        bufs.append(p)
        return p

    try:
        for L, seed, sigma in ((48, 1, 0.5), (24, 2, 0.7)):  # the second call is smaller
            print(f"[no-alloc] L={L} sigma={sigma}", flush=True)
            Mnum, Wt = sums(L, seed)
            M_ref, avg_ref = existing(L, Mnum, Wt, sigma)
            n = L ** 3
            dM, dW, dmodel, davg = dev_alloc(4 * n), dev_alloc(4 * n), dev_alloc(4 * n), dev_alloc(8)
            M = np.full(n, 7.0, np.float32)  # the device model starts as something else
            assert hip.hipMemcpy(dM, Mnum.ctypes.data, 4 * n, H2D) == 0
            assert hip.hipMemcpy(dW, Wt.ctypes.data, 4 * n, H2D) == 0
            assert hip.hipMemcpy(dmodel, M.ctypes.data, 4 * n, H2D) == 0
            assert hip.hipDeviceSynchronize() == 0
            check(lib().emc_compress_device(L, sigma, dM, dW, dmodel, davg, None))
            assert hip.hipDeviceSynchronize() == 0
            avg_h = np.zeros(1, np.float64)
            assert hip.hipMemcpy(M.ctypes.data, dmodel, 4 * n, D2H) == 0
            assert hip.hipMemcpy(avg_h.ctypes.data, davg, 8, D2H) == 0
            avg = float(avg_h[0])
            rel = float(np.linalg.norm(M - M_ref) / np.linalg.norm(M_ref))
            print(f"   model rel L2 vs emc_compress {rel:.2e}, equal {np.array_equal(M, M_ref)}, "
                  f"avg {avg:.9g} / {avg_ref:.9g}", flush=True)
            assert avg == avg_ref                   # same kernels, same order: the same bits
            assert np.array_equal(M, M_ref)
    finally:
        for p in bufs:
            hip.hipFree(p)


def test_compress_device_rejects_sizes_beyond_its_scratch():
    with pytest.raises(EmcError, match="outside 4..128"):
        check(lib().emc_compress_device(129, 0.5, C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), None, None))
    with pytest.raises(EmcError):
        check(lib().emc_compress_device(32, 0.5, None, C.c_void_p(16), C.c_void_p(16), None, None))
