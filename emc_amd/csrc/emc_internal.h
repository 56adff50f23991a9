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
// Partial sums of the model statistics (256 blocks x 4 values).
constexpr int kRedLen = 1024;

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
    SC_LOSS = 11,     // proposal network: summed cross-entropy of a training pass
    SC_NNZ = 12,      // non-zero responsibilities of the last dense E-step (BF16 hi or lo non-zero)
    SC_SKIPPED = 13,  // (band, pattern) logit blocks whose stores the logL kernel left out
    SC_N = 16
};

// Learned orientation proposal: network, training state and the grouped (per coarse cell) work
// lists of a proposal iteration.
struct Proposal {
    emc_proposal_options o{};
    int C = 0, Cp = 0, H = 0;           // cells, cells padded to 64, hidden width
    static constexpr int D = 1024;      // features: 32 x 32 pooled pattern
    int iter = 0;                       // iterations run since enable
    int passes = 0;                     // training passes run
    int adam_t = 0;                     // Adam step counter
    double last_loss = -1.0;
    // partition of the fine rotations
    std::vector<int32_t> cell_h, rbeg_h, cell_rot_h;  // [S], [C+1], [S]
    int32_t *cell_of = nullptr, *cell_rot = nullptr, *rbeg = nullptr;
    // parameters: one FP32 vector W1 | b1 | W2 | b2 | W3 [Cp][H] | b3 [Cp] (rows >= C zero)
    int64_t n_par = 0, oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
    float *par = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
    uint16_t *par16 = nullptr;
    // features
    float *mu = nullptr, *istd = nullptr;  // [D]
    uint16_t *X = nullptr;                 // [Ip][D] standardised features of the patterns
    // activations, rows_cap rows (inference runs over all patterns at once)
    int rows_cap = 0;
    float *pre1 = nullptr, *pre2 = nullptr, *z = nullptr, *prob = nullptr;
    uint16_t *h1 = nullptr, *h2 = nullptr, *dz = nullptr, *d1 = nullptr, *d2 = nullptr;
    double *loss_part = nullptr;
    // training
    float *Tt = nullptr;           // [Cp][Ip] cell sums of the responsibilities
    float *Mx = nullptr;           // [Cp][Ip] largest logit per cell (when recall > 0)
    bool targets_valid = false;
    uint16_t *Xb = nullptr;        // [batch][D]
    float *tb = nullptr;           // [batch][Cp]
    int32_t *bidx = nullptr;       // [I + n_syn] shuffled training indices
    int n_syn = 0;
    float *syn_counts = nullptr;   // [n_syn][Pp]
    uint16_t *Xs = nullptr;        // [n_syn][D]
    int32_t *syn_rot = nullptr;    // [n_syn]
    uint64_t rng_state = 0;
    // prediction
    int32_t *topc = nullptr;       // [Ip][top_k]
    float *pmass = nullptr;        // [Ip]
    double *cap = nullptr;         // [Ip] captured mass per pattern
    int32_t *topc_h = nullptr;     // pinned host copies of the last prediction: [I][top_k]
    float *pmass_h = nullptr;      // [I]
    hipEvent_t ev_pred = nullptr;  // recorded after the prediction's copies to the host
    bool pred_valid = false;
    std::vector<double> cap_h;     // per-pattern captured mass of the last measurement
    // recall: cells of largest max logit in each pattern's most recent search, -1 = none
    int32_t *recall = nullptr;     // [Ip][recall]
    int32_t *recall_h = nullptr;   // [I][recall] pinned host copy taken with each prediction
    // grouped work lists (built on the host from the prediction)
    int64_t n_slots = 0, n_blk = 0;  // slots, block elements
    int n_fallback = 0;
    double work_fraction = 0.0;
    int n_tiles_e = 0, n_tiles_m = 0, n_tiles_sm = 0;
    int64_t list_cap = 0;          // int32 elements of `lists`
    int32_t *lists = nullptr;      // device copy of the packed lists below
    int32_t *lists_h = nullptr;    // pinned host staging
    int64_t *boff = nullptr;       // [C] element offset of each cell's logits / P block
    int64_t *boff_h = nullptr;     // pinned host staging
    float *raw = nullptr;          // [max(Ip, n_syn)][D] unstandardised features (scratch)
    // offsets (in int32 elements) of each list inside `lists`
    int64_t off_slot_pat = 0, off_sbeg = 0, off_nsp = 0, off_pbeg = 0, off_pslot = 0, off_tile_e = 0,
            off_tile_m = 0, off_tile_sm = 0;
    int64_t slot_cap = 0;          // doubles of slot scratch
    double *slot_ps = nullptr;     // [slots] partial sums
    float *slot_pm = nullptr;      // [slots] partial maxima
    double *slot_lse = nullptr;    // [slots]
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
    bool chi2_mfma = false;          // a chi^2 diff on the MFMA path (expanded contractions)
    double sumK_total = 0.0;         // sum of all counts over active pixels (FP64)
    int32_t max_count = 0;
    std::string err;

    bool has_rot = false, has_coords = false, has_mask = false, has_K = false, has_model = false;
    bool has_logl = false, has_resp = false, has_mstep = false;
    bool has_lse = false;  // lse and T of the last E-step (full or grouped) are current

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
    float4 *modelq = nullptr; // [N3] quad-packed model of the trilinear statistics pass (MFMA path)
    float *taps = nullptr;    // [L] blur taps
    // slicing stage
    double *sumW = nullptr;   // [Sp]
    int32_t *ns = nullptr;    // [Sp]
    double *colsum = nullptr; // [Pp][2] per pixel: sum over rotations of log W, valid count
    float2 *colp = nullptr;   // [ceil(S / 16)][Pp] per statistics block: (sum of log W, valid count)
    double2 *colpart = nullptr; // [32][Pp] fixed-order partial column sums
    float *colmean = nullptr; // [Pp]
    float *bias = nullptr;    // [Sp] per-rotation additive term of the logits
    float *rscale = nullptr;  // [Sp] chi^2 factor 1/(2 n_s sigma^2)
    double *cst = nullptr;    // [Ip] per-pattern constant removed from the stored logits
    float *A0 = nullptr, *A1 = nullptr;                            // [chunkp][Pp] FP32 slices
    // intensity scaling (options.scaling): the two operands of its sums, counts squared,
    // 1/phi per (rotation, pattern) and the M-step's operand P/phi
    float *Am = nullptr, *AWm = nullptr;  // [chunkp][Pp] 1[W > 1e-10], W 1[W > 1e-10]
    float *K2 = nullptr;                  // [Ip][Pp] K^2
    float *iphi = nullptr;                // [Sp][Ip] (1 until the first likelihood pass)
    float *Pphi = nullptr;                // [Sp][Ip]
    // chi^2 kinds on the MFMA path: the BF16-split operands of the four contractions
    // (a W, a, 1[W > 1e-10], W 1[W > 1e-10]; la_terms terms each), counts squared as an exact
    // hi + lo BF16 pair, per-rotation sum a W^2 and the products' FP32 buffers
    emc::bf16_t *Cx = nullptr;                 // [4][la_terms][chunkp][Pp]
    emc::bf16_t *K2h = nullptr, *K2l = nullptr; // [Ip][Pp]
    bool k2_split = false;                     // some count above 16: K^2 needs the lo term
    double *Aw = nullptr;                      // [chunkp]
    float *G2 = nullptr, *G3 = nullptr, *G4 = nullptr; // [chunkp][Ip]
    emc::bf16_t *Lh = nullptr, *Lm = nullptr, *Ll = nullptr;       // [chunkp][Pp] BF16 log W terms
    float *lw32 = nullptr;    // [Sp][Pp] log W from the statistics pass (-inf where W <= 0), MFMA path
    // responsibilities
    float *R = nullptr;                                // [Sp][Ip] logits (P in place on FP32 path)
    emc::bf16_t *Ph = nullptr, *Pl = nullptr;          // [Sp][Ip] BF16 hi/lo P (MFMA path)
    uint8_t *pdirty = nullptr;  // [Sp][ceil(I/256)]: 0 = that 256-pattern segment of Ph/Pl holds only zeros
    int32_t *runmax = nullptr;  // [Ip] per pattern, the largest band maximum of the logits found so far in
                                // this logL pass (order-preserving integer form of a float)
    int16_t *blk = nullptr;     // [fused bands][Ip] where the skipping logL pass put each band's logits of
                                // a pattern: block row (>= 0), -2 in place, -1 not stored
    bool skip_logits = false;   // the current logL pass may leave out negligible bands' logit stores
    bool logits_partial = false; // the last logL pass left some out
    double *lse = nullptr;                             // [Ip] lse' (constants excluded)
    double *part = nullptr;                            // softmax / reduction partials
    double *red = nullptr;                             // [kRedLen] model-statistics partials (own buffer:
                                                       // part may hold the logL epilogue's softmax partials)
    int64_t part_len = 0;
    double *T = nullptr;                               // [Sp]
    float *U = nullptr;                                // [chunkp][Pp] (atomic insertion)
    double *scal = nullptr;                            // [SC_N] device scalars
    double *scal_h = nullptr;                          // pinned host mirror
    // binned insertion: valid (s, j) pairs sorted by (chunk, model brick), cut into units of at most
    // 4,096 pairs; per unit a base-voxel permutation; partial halo bricks per block (<= 4 units of a
    // brick, one workgroup) and per group of blocks
    uint32_t *bin_vals = nullptr;                      // [S * Pa] packed s << jbits | j, (s, j) order per brick
    uint16_t *bin_perm = nullptr;                      // [S * Pa] each pair's position in base-voxel order in its unit
    uint16_t *bin_vbeg = nullptr;                      // [n_units][520] start of each base voxel's run
    float *bin_frac = nullptr;                         // [n_units][3][4096] trilinear fractions (base-voxel order)
    int32_t *bin_units = nullptr;                      // [n_units][4] = key, begin, end, 0
    int32_t *bin_blocks = nullptr;                     // [n_blocks][4] = key, first unit, end unit, 0 (one workgroup)
    int32_t *bin_groups = nullptr;                     // [n_groups][4] = key, first block, end block (per chunk), 0
    int32_t *bin_gbeg = nullptr;                       // [chunks][bricks + 1] first group of each brick
    float *bin_part = nullptr;                         // [blocks of a chunk][2][9^3] block partials
    float *bin_gpart = nullptr;                        // [groups of a chunk][2][9^3] group partials
    std::vector<int> bin_chunk_blocks, bin_chunk_groups; // block / group range of each rotation chunk
    bool bins_valid = false;
    // learned proposal (null when disabled)
    emc::Proposal *prop = nullptr;
};

namespace emc {

// ---- launchers (each enqueues on `st`, returns hipError_t) --------------------------------
hipError_t launch_rotmat(const float *q, const float *w, int S, int Sp, float *rotm, float *logw,
                         hipStream_t st);
hipError_t launch_fill_f32(float *p, float v, int64_t n, hipStream_t st);
hipError_t launch_fill_f64(double *p, double v, int64_t n, hipStream_t st);
hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t st);

// Slicing statistics over rotations [0, S): sumW, ns and (if colsum) sum of log W per pixel.
// colp / colpart hold colp_len / colpart_len elements.
int64_t colp_len(int S, int Pp);
int64_t colpart_len(int Pp);
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

// FP32 contractions (plain path).  With options.scaling the logits use iphi (launch_scaling_f32
// first) and the M-step reads P/phi from iphi (launch_pscale_f32 after the responsibilities).
hipError_t launch_logl_f32(const emc_ctx *c, int s0, int cs, hipStream_t st);
hipError_t launch_mstep_f32(const emc_ctx *c, int s0, int cs, hipStream_t st);
// 1/phi of rotations [s0, s0+cs) from the chunk's FP32 slices: sum K^2 and sum K W over
// {W > 1e-10} as two contractions over pixels, then the quotient (1 where either sum is 0).
hipError_t launch_scaling_f32(const emc_ctx *c, int s0, int cs, hipStream_t st);
// Pphi = P * iphi for all rotations (P in R).
hipError_t launch_pscale_f32(const emc_ctx *c, hipStream_t st);
// K2 = K32^2 (after the counts are set).
hipError_t launch_square_counts(const emc_ctx *c, hipStream_t st);
// MFMA contractions.
hipError_t launch_logl_mfma(const emc_ctx *c, int s0, int cs, hipStream_t st);
// C [chunkp][Ip] (=, or += with accumulate) sum over pixels of A [chunkp][Pp] (na BF16 terms) and
// the counts-like operand B [Ip][Pp] (BF16, or bytes when B8 is given): the large-tile kernel.
hipError_t launch_gemm_store_mfma(const emc_ctx *c, int na, const bf16_t *A0, const bf16_t *A1, const bf16_t *A2,
                                  const bf16_t *B16, const uint8_t *B8, float *C, bool accumulate, hipStream_t st);
// chi^2 logits of rotations [s0, s0+cs) on the MFMA path, from the chunk's FP32 slices (A0 = W,
// A1 = the pixel weight a): D = sum a W^2 - 2 ip sum a W K + ip^2 sum a K^2 over the valid pixels,
// ip = 1/phi (options.scaling: sum W K / sum K^2 over W > 1e-10; else 1); R = log w_s - D rscale_s.
// With scaling, iphi rows [s0, s0+cs) are written for the M-step.
hipError_t launch_logl_chi2_mfma(const emc_ctx *c, int s0, int cs, hipStream_t st);
// Ph / Pl segments of 256 patterns tracked by pdirty, and marking every one as possibly non-zero
// (after anything but the dense softmax wrote Ph / Pl).
inline int p_segments(const emc_ctx *c) { return (c->I + 255) / 256; }
hipError_t launch_mark_p_dirty(const emc_ctx *c, hipStream_t st);

// The dense M-step returns at once when gate is non-null and *gate <= gate_thr.
hipError_t launch_mstep_mfma(const emc_ctx *c, int s0, int cs, bool fused, const double *gate, double gate_thr,
                             hipStream_t st);
// Sparse M-step: U rows [s0, s0+cs) summed over the non-zero responsibilities of each rotation
// (same U as the dense contraction up to summation order); returns at once when gate is non-null
// and *gate > gate_thr.  Needs Pp <= kSparseMaxPp.
constexpr int kSparseMaxPp = 16384;
hipError_t launch_mstep_sparse(const emc_ctx *c, int s0, int cs, const double *gate, double gate_thr, hipStream_t st);

// Responsibilities.  When logl_fuses_softmax(c), the logL contraction's epilogue writes the
// per-band softmax partials (one band per 256 rotation rows) and launch_softmax skips its pass.
bool logl_fuses_softmax(const emc_ctx *c);
int softmax_fused_bands(const emc_ctx *c);
// The logL pass leaves out the stores of negligible bands' logits (c->skip_logits on the DMA kernel).
bool logl_skips_stores(const emc_ctx *c);
int64_t softmax_part_len(int S, int I, int fused_bands);
hipError_t launch_softmax(const emc_ctx *c, hipStream_t st);
// Atomic insertion of U rows [s0, s0+cs).
hipError_t launch_insert(const emc_ctx *c, int s0, int cs, hipStream_t st);
// Brick-binned insertion: build the sorted pair lists (setup; allocates), then per chunk.
hipError_t build_bins(emc_ctx *c, std::string &err);
hipError_t launch_insert_binned(const emc_ctx *c, int s0, int cs, hipStream_t st);
// Model statistics (max, sum of Wt) and the normaliser: the mean of the model over the voxels
// with weight (support: Wt > 0, after the divide) or over the positive voxels; apply scales every
// voxel by its inverse.
hipError_t launch_normalize(const emc_ctx *c, bool apply, bool support, hipStream_t st);
// Blur taps of side L (FP64 sums, stored as float).
hipError_t launch_blur_taps(int L, double sigma, float *taps, hipStream_t st);
// Compress on explicit buffers: divide, normalise over the voxels with weight (FP64 partials in
// red [kRedLen], scalars in scal [SC_N]), blur when asked (t1, t2: two L^3 volumes; t2 may be
// Mnum, which the divide has consumed).
hipError_t launch_compress_buffers(int L, bool blur, const float *Mnum, const float *Wt, float *model,
                                   const float *taps, float *t1, float *t2, double *red, double *scal, hipStream_t st);
// Deterministic single-workgroup sum of n doubles into *out.
hipError_t launch_sum_f64(const double *x, int n, double *out, hipStream_t st);

// ---- learned proposal ------------------------------------------------------------------------
// Grouped (per coarse cell) contractions and softmax over the work lists in c->prop.
hipError_t launch_grouped_logl(const emc_ctx *c, hipStream_t st);
hipError_t launch_grouped_softmax(const emc_ctx *c, hipStream_t st);
hipError_t launch_grouped_mstep(const emc_ctx *c, hipStream_t st);
// Recalled cells: each pattern's cells of largest max logit among the searched slots (after
// launch_grouped_softmax) or among all cells (Mx, after a full search).
hipError_t launch_recall_grouped(const emc_ctx *c, hipStream_t st);
hipError_t launch_recall_dense(const emc_ctx *c, hipStream_t st);
constexpr int kMaxRecall = 8;  // largest number of recalled cells per pattern

// Network kernels.  mlp_gemm: C[m,n] = sum_k A(m,k) B(n,k), A/B "col" = stored [k][m] / [k][n];
// M, N multiples of 64, K of 32.
enum { MLP_EPI_STORE = 0, MLP_EPI_BIAS_GELU = 1, MLP_EPI_BIAS = 2, MLP_EPI_GELU_GRAD = 3 };
hipError_t mlp_gemm(bool acol, bool bcol, int epi, const bf16_t *A, int lda, const bf16_t *B, int ldb, int M, int N,
                    int K, float *out32, bf16_t *out16, int ldo, const float *bias, const float *pre, int m_valid,
                    int n_valid, hipStream_t st);
hipError_t mlp_features(const float *K32, const bf16_t *K16, const uint8_t *K8, int ld, int Pa, const int32_t *pix,
                        int L, int rows, int rows_pad, float *raw, hipStream_t st);
hipError_t mlp_feat_stats(const float *raw, int rows, float *mu, float *istd, hipStream_t st);
hipError_t mlp_standardise(const float *raw, int rows, int rows_pad, const float *mu, const float *istd, bf16_t *X,
                           hipStream_t st);
hipError_t mlp_softmax_ce(const float *z, int Cp, int C, int rows_valid, int rows, const float *t, float scale,
                          float *p, bf16_t *dz, double *loss_part, hipStream_t st);
hipError_t mlp_accum(const double *x, int n, double *acc, hipStream_t st);
hipError_t mlp_colsum(const bf16_t *d, int rows, int ld, int n_valid, float *g, hipStream_t st);
hipError_t mlp_adam(float *w, const float *g, float *m, float *v, bf16_t *w16, int64_t n, float lr, int t,
                    hipStream_t st);
hipError_t mlp_to_bf16(const float *w, int64_t n, bf16_t *w16, hipStream_t st);
hipError_t mlp_topk(const float *p, int Cp, int C, int rows, int k, int32_t *cells, float *mass, hipStream_t st);
// Cell sums Tt of the full responsibilities; with Mx also each cell's largest logit (reads the
// logits c->R, so only while they are resident: right after a full-search E-step).
hipError_t mlp_cell_targets(const emc_ctx *c, float *Tt, float *Mx, hipStream_t st);
hipError_t mlp_captured(const emc_ctx *c, double *cap, hipStream_t st);
// Adds the recalled cells' predicted probability (those not in the top k) to pmass.
hipError_t mlp_union_mass(const emc_ctx *c, hipStream_t st);
hipError_t mlp_gather_batch(const emc_ctx *c, const int32_t *idx, int rows, hipStream_t st);
hipError_t mlp_synth_counts(const emc_ctx *c, int n, float photons, uint64_t seed, hipStream_t st);

// Host side of the proposal (proposal.cpp); each sets c->err on failure.
void proposal_free(Proposal *p);
// Network prediction for every pattern into topc / pmass (device) and topc_h / pmass_h /
// recall_h (host): enqueue, then wait for the host copies (the stream may run other work
// enqueued in between).
emc_status proposal_predict_enqueue(emc_ctx *c);
emc_status proposal_predict_wait(emc_ctx *c);
emc_status proposal_predict(emc_ctx *c);
// Grouped work lists from the prediction; *too_many = the fallback share exceeds max_fallback.
emc_status proposal_build_lists(emc_ctx *c, bool *too_many);
// Cell targets from the dense responsibilities; captured mass of the last prediction if asked.
emc_status proposal_targets(emc_ctx *c, bool measure, bool logits, double *captured);
// Training passes over the patterns' cell targets (use_patterns) and syn_fraction of the
// synthetic patterns, sliced afresh from the current model each pass; loss of the last pass.
emc_status proposal_train(emc_ctx *c, int epochs, bool use_patterns, double syn_fraction, double *loss);

// BF16 helpers (host and device agree: round to nearest even, finite inputs).
__host__ __device__ inline bf16_t f2bf(float x)
{
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }
// The same rounding on the gfx950 conversion instruction (one v_cvt_pk_bf16_f32 per value; equal to
// f2bf for every non-NaN float, checked over all 2^32 bit patterns on the device).  Scalar form only:
// the two-element vector conversion, __builtin_convertvector(float2 -> __bf16 x 2), is lowered to one
// instruction that converts the first element alone.
__device__ inline bf16_t f2bf_dev(float x) { return __builtin_bit_cast(bf16_t, (__bf16)x); }

}  // namespace emc
