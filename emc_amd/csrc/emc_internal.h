// Internal declarations of libemc: the context and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/emc.h"

namespace emc {

using bf16_t = uint16_t;  // BF16 storage (bit pattern); MFMA operands are bitcast from it

// Tile sizes the device buffers are padded to (rows of R / P, patterns, compacted pixels).
constexpr int kPadS = 256;
constexpr int kPadI = 256;
constexpr int kPadP = 128;

inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// Device scalars block (ctx->scal): indices
enum {
    SC_PHI0 = 0,      // global fluence used this iteration
    SC_LOGPHI0 = 1,   // log(phi0)
    SC_TOTAL = 2,     // sum_i lse'_i (without the per-pattern constants)
    SC_AVG = 3,       // model average before normalisation
    SC_INV = 4,       // 1/avg applied
    SC_COUNT = 5,     // positive voxel count
    SC_MAX = 6,       // model max
    SC_WSUM = 7,      // sum of Wt
    SC_NONFINITE = 8, // > 0 if a non-finite lse or normaliser appeared
    SC_SUMW = 9,      // sum over rotations of the slice sums
    SC_B0 = 10,       // global logit shift removed from the stored logits (true Poisson)
    SC_N = 16
};

}  // namespace emc

struct emc_ctx {
    emc_dims dims{};
    emc_options opt{};
    int L = 0, S = 0, I = 0, P = 0;  // side, rotations, patterns, pixels (L*L)
    int64_t N3 = 0;                  // L^3
    int Pa = 0, Pp = 0;              // active (mask != 0) pixels, padded count
    int Sp = 0, Ip = 0;              // padded rotations / patterns
    int chunk = 0, chunkp = 0;       // rotations per chunk, padded
    bool mfma = false;               // precision is a BF16 MFMA mode
    int la_terms = 2;                // BF16 terms of log W (2 or 3)
    bool k_split = false;            // counts need a hi+lo BF16 pair (max count > 256)
    double sumK_total = 0.0;         // sum of all counts over active pixels (FP64)
    int32_t max_count = 0;
    std::string err;

    bool has_rot = false, has_coords = false, has_mask = false, has_K = false, has_model = false;
    bool has_logl = false, has_resp = false, has_mstep = false;

    hipStream_t stream = nullptr;
    hipEvent_t ev[2 * EMC_N_STAGES] = {};  // start/stop per stage
    std::vector<hipEvent_t> gev;           // start/stop per chunk: logL and M-step contractions

    // geometry and inputs
    float *q = nullptr;       // [S][4]
    float *rotm = nullptr;    // [Sp][12] rotation matrices (rows >= S zero)
    float *logw = nullptr;    // [Sp]
    float *coords = nullptr;  // [3][P] full detector coordinates
    int32_t *mask = nullptr;  // [P]
    float *axyz = nullptr;    // [3][Pp] coordinates of active pixels (padding 0)
    int32_t *pix = nullptr;   // [Pp] detector index of active pixel j (-1 padding)
    float *K32 = nullptr;     // [Ip][Pp] counts, FP32 path
    float *KT32 = nullptr;    // [Pp][Ip]
    emc::bf16_t *K16 = nullptr, *K16lo = nullptr;    // [Ip][Pp] MFMA path
    emc::bf16_t *KT16 = nullptr, *KT16lo = nullptr;  // [Pp][Ip]
    uint8_t *K8 = nullptr, *KT8 = nullptr;            // byte copies when every count <= 255
    double *sumK = nullptr;   // [Ip]
    // model
    float *model = nullptr, *Mnum = nullptr, *Wt = nullptr, *tmp3 = nullptr;  // [N3]
    float *taps = nullptr;    // [L] blur taps
    // slicing stage
    double *sumW = nullptr;   // [Sp]
    int32_t *ns = nullptr;    // [Sp]
    double *colsum = nullptr; // [Pp][2] per pixel: sum over rotations of log W, valid count
    float *colmean = nullptr; // [Pp]
    float *bias = nullptr;    // [Sp] per-rotation additive term of the logits
    float *rscale = nullptr;  // [Sp] chi^2 factor 1/(2 n_s sigma^2)
    double *cst = nullptr;    // [Ip] per-pattern constant removed from the stored logits
    float *A0 = nullptr, *A1 = nullptr;                            // [chunkp][Pp] FP32 slices
    emc::bf16_t *Lh = nullptr, *Lm = nullptr, *Ll = nullptr;       // [chunkp][Pp] BF16 log W terms
    // responsibilities
    float *R = nullptr;                                // [Sp][Ip] logits (P in place on FP32 path)
    emc::bf16_t *Ph = nullptr, *Pl = nullptr;          // [Sp][Ip] BF16 hi/lo P (MFMA path)
    double *lse = nullptr;                             // [Ip] lse' (constants excluded)
    double *part = nullptr;                            // softmax / reduction partials
    int64_t part_len = 0;
    double *T = nullptr;                               // [Sp]
    float *U = nullptr;                                // [chunkp][Pp] (atomic insertion)
    double *scal = nullptr;                            // [SC_N] device scalars
    double *scal_h = nullptr;                          // pinned host mirror
    // binned insertion: (s, j) pair indices sorted by (chunk, model brick); work units
    uint32_t *bin_vals = nullptr;                      // [S * Pa]
    int32_t *bin_units = nullptr;                      // [n_units][3] = key, begin, end
    std::vector<int> bin_chunk_units;                  // unit range of each rotation chunk
    bool bins_valid = false;
};

namespace emc {

// ---- launchers (each enqueues on `st`, returns hipError_t) --------------------------------
hipError_t launch_rotmat(const float *q, const float *w, int S, int Sp, float *rotm, float *logw,
                         hipStream_t st);
hipError_t launch_fill_f32(float *p, float v, int64_t n, hipStream_t st);
hipError_t launch_fill_f64(double *p, double v, int64_t n, hipStream_t st);

// Slicing statistics over rotations [0, S): sumW, ns and (if colsum) sum of log W per pixel.
hipError_t launch_slice_stats(const emc_ctx *c, bool want_colsum, hipStream_t st);
// Fluence, per-rotation bias / chi^2 scale, column means.
hipError_t launch_bias(const emc_ctx *c, float sigma, hipStream_t st);
// Per-pattern constant: cst[i] = sum_j K[i,j] colmean[j] (MFMA path) - B0.
hipError_t launch_pattern_const(const emc_ctx *c, bool centred, hipStream_t st);
// FP32 slices of rotations [s0, s0+cs) into A0/A1 (rows up to chunkp zero-padded).
hipError_t launch_slice_f32(const emc_ctx *c, int s0, int cs, hipStream_t st);
// BF16-split centred log-slices into Lh/Lm/Ll.
hipError_t launch_slice_bf16(const emc_ctx *c, int s0, int cs, hipStream_t st);
// Full-detector slice values of rotations [s0, s0+cs) into out [cs][P] (-1 invalid/masked).
hipError_t launch_expand(const emc_ctx *c, int s0, int cs, float *out, hipStream_t st);

// FP32 contractions (plain path).
hipError_t launch_logl_f32(const emc_ctx *c, int s0, int cs, hipStream_t st);
hipError_t launch_mstep_f32(const emc_ctx *c, int s0, int cs, hipStream_t st);
// MFMA contractions.
hipError_t launch_logl_mfma(const emc_ctx *c, int s0, int cs, hipStream_t st);
hipError_t launch_mstep_mfma(const emc_ctx *c, int s0, int cs, bool fused, hipStream_t st);

// Responsibilities.
hipError_t launch_softmax(const emc_ctx *c, hipStream_t st);
// Atomic insertion of U rows [s0, s0+cs).
hipError_t launch_insert(const emc_ctx *c, int s0, int cs, hipStream_t st);
// Brick-binned insertion: build the sorted pair lists (setup; allocates), then per chunk.
hipError_t build_bins(emc_ctx *c, std::string &err);
hipError_t launch_insert_binned(const emc_ctx *c, int s0, int cs, hipStream_t st);
// Compress: divide, blur, normalise.
hipError_t launch_divide(const emc_ctx *c, hipStream_t st);
hipError_t launch_blur(const emc_ctx *c, hipStream_t st);
hipError_t launch_normalize(const emc_ctx *c, bool apply, hipStream_t st);

// BF16 helpers (host and device agree: round to nearest even, finite inputs).
__host__ __device__ inline bf16_t f2bf(float x)
{
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

}  // namespace emc
