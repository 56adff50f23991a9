/*
 * emc.h -- C ABI of the MI355X-native EMC engine (libemc.so).
 *
 * One context binds one HIP device and one stream and owns every device buffer.  Host pointers
 * passed in are borrowed for the duration of the call only.  Every call returns an emc_status;
 * emc_last_error() gives the message.  Nothing prints and nothing exits.
 *
 * Each entry point names the reference interface it replaces (JinLi971/EMC, files at the top
 * level of that repository).  The reference exposes C++ free functions over raw device
 * pointers (emc_cuda_help.h:9-117) driven by main.cpp:239-516; here the same stage granularity
 * is kept (so each stage is testable on its own) but buffers live inside the context.
 *
 * Layouts at this boundary are the reference's: pixel index p = y*L + x; model index
 * (z*L + y)*L + x; rotations [S][4] (q0,q1,q2,q3); weights [S]; patterns [I][L*L];
 * responsibilities and logits [S][I] (row = rotation).
 */
#ifndef EMC_H
#define EMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum emc_status {
    EMC_OK = 0,
    EMC_E_INVALID = 1,     /* bad argument or shape */
    EMC_E_HIP = 2,         /* HIP runtime error */
    EMC_E_OOM = 3,         /* device or host allocation failed */
    EMC_E_NONFINITE = 4,   /* a non-finite value appeared where it must not (lse, sums) */
    EMC_E_UNSUPPORTED = 5, /* valid request this build does not support */
    EMC_E_STATE = 6,       /* call order violated (e.g. logl before set_model) */
    EMC_E_IO = 7           /* file could not be read or written */
} emc_status;

/* configs.h:11  enum diff_type {absolute=0, poisson=1, relative=2}.  The reference's "poisson"
 * is the chi^2 form sum (W-K)^2/(W+0.02) (emc_cuda.cu:196-216).  EMC_DIFF_TRUE_POISSON is the
 * Poisson log-likelihood K.log W - sum W with one global fluence (no per-pattern scaling). */
enum { EMC_DIFF_ABSOLUTE = 0, EMC_DIFF_POISSON = 1, EMC_DIFF_RELATIVE = 2, EMC_DIFF_TRUE_POISSON = 3 };

/* Slicing/insertion interpolation.  NEAREST reproduces lroundf and the reference's bounds
 * (emc_cuda.cu:129-134, emc_atomic.cu:62-69); TRILINEAR uses the 8 corners of [0, L-1]^3. */
enum { EMC_INTERP_NEAREST = 0, EMC_INTERP_TRILINEAR = 1 };

/* Arithmetic of the two contractions (logL and the M-step sum).
 *   FP32   : direct FP32 (vector ALU), the plain path.
 *   BF16X2 : MFMA, counts exact in BF16, log W and P each split into BF16 hi+lo, FP32 accumulate.
 *   BF16X3 : as BF16X2 with log W split into three BF16 terms (hi+mid+lo).
 * Counts above 256 and non-integer patterns are split into two BF16 terms on the MFMA modes.
 * The chi^2 kinds run on MFMA as expanded contractions (sum a W^2 - 2 sum a W K + sum a K^2, with
 * K^2 as two BF16 terms, exact for integer counts <= 255). */
enum { EMC_PREC_FP32 = 0, EMC_PREC_BF16X2 = 1, EMC_PREC_BF16X3 = 2 };

/* How the M-step inserts the weighted slice sums into the model.
 *   ATOMIC : U = P^T.K is materialised per rotation chunk, then scattered with float atomics
 *            (emc_atomic.cu:33-85, 183-200).
 *   FUSED  : the MFMA M-step scatters from its accumulators with float atomics (no U buffer).
 *   BINNED : U is materialised; (rotation, pixel) pairs pre-sorted by model brick are
 *            accumulated per brick in LDS and flushed once (no contended global atomics). */
enum { EMC_INSERT_ATOMIC = 0, EMC_INSERT_FUSED = 1, EMC_INSERT_BINNED = 2 };

/* How the MFMA path forms U = P^T.K (emc_options.mstep).  SPARSE sums, per rotation, only the
 * patterns whose responsibility is non-zero (the same sums up to their order); AUTO uses it when
 * at most 2 % of the responsibilities are non-zero.  FUSED insertion always uses DENSE. */
enum { EMC_MSTEP_AUTO = 0, EMC_MSTEP_DENSE = 1, EMC_MSTEP_SPARSE = 2 };

typedef struct emc_dims {
    int32_t model_side;  /* L: model L^3, patterns L x L (configs.h:14 model_side) */
    int32_t n_rotations; /* S (rotations.c:457 return value) */
    int32_t n_patterns;  /* I (configs.h:22 N_images) */
    int32_t slice_chunk; /* rotations per chunk, 0 = all (configs.h:21 slice_chunk) */
} emc_dims;

typedef struct emc_options {
    int32_t diff;            /* EMC_DIFF_*   (configs.h:52 diff) */
    int32_t interp;          /* EMC_INTERP_* */
    int32_t precision;       /* EMC_PREC_*   */
    int32_t insert;          /* EMC_INSERT_* */
    double sigma_start;      /* configs.h:47 */
    double sigma_final;      /* configs.h:48 */
    int32_t sigma_half_life; /* configs.h:49 */
    int32_t device;          /* HIP device ordinal */
    double blur_sigma;       /* model blur (voxels) after divide-by-weight; 0 = off (configs.h:80) */
    double fluence;          /* TRUE_POISSON photons per model unit; 0 = automatic per iteration */
    int32_t timing;          /* 1 = record per-stage HIP events (emc_iter_stats.stage_ms) */
    int32_t mfma_tile;       /* MFMA contraction tiling: 0 = 256-row tiles (default), 1 = 128x128 */
    int32_t count_bytes;     /* counts operand of the 256-row MFMA kernels: 0 = one byte per count when
                                every count <= 255 (exact; widened to BF16 in registers), 2 = BF16 */
    int32_t mstep;           /* MFMA M-step sum P^T.K: EMC_MSTEP_AUTO (default) chooses on the device,
                                from the E-step's count of non-zero responsibilities, between the
                                dense contraction and the sparse one (only the non-zero P); or
                                EMC_MSTEP_DENSE / EMC_MSTEP_SPARSE */
    int32_t scaling;         /* 1 = per-(rotation, pattern) intensity scaling, the reference's
                                known_intensity = false (main.cpp:283-295): phi[s,i] =
                                sum K^2 / sum K W over {mask, W > 1e-10, K > 1e-10}
                                (emc_cuda.cu:438-454, 516-532), the chi^2 kinds compare W with K/phi
                                and the M-step sums K P / phi (emc_atomic.cu:106-108); both sums
                                of phi are contractions over pixels.  chi^2 kinds only (any
                                precision).  0 = phi = 1 (known_intensity = true) */
    int32_t tile_rows;       /* large-tile MFMA contractions: row panels per group of consecutive tiles
                                (L2 reuse within an XCD); 0 = the built-in default */
    int32_t mfma_staging;    /* 256-row MFMA contractions of a two-term operand with a one-term one
                                (logL, dense M-step, chi^2 sums): 0 = operands staged in LDS by direct
                                global->LDS loads (BF16 counts), 1 = staged through registers (byte
                                counts widened in registers).  Same products in the same order:
                                identical results */
    int32_t keep_logits;     /* emc_iterate with the MFMA likelihood: 0 = the logL kernel leaves out
                                the stores of a 256-rotation band's logits for a pattern when the
                                band's largest logit lies more than 128 below a largest logit
                                already found for that pattern (its responsibilities there are
                                exactly 0; the softmax reads those logits as -inf).  The iteration's
                                results are the same; the logits are then incomplete and
                                emc_get_logl after emc_iterate fails.  1 = every logit is stored.
                                emc_logl always stores every logit, and the learned proposal's
                                recall (which reads them) keeps them too */
} emc_options;

/* Slots of emc_iter_stats.stage_ms.  LOGL_GEMM / MSTEP_GEMM are the contraction kernels alone
 * (summed over rotation chunks); the other slots bracket whole stages. */
enum {
    EMC_STAGE_SLICE = 0,
    EMC_STAGE_LOGL = 1,
    EMC_STAGE_SOFTMAX = 2,
    EMC_STAGE_MSTEP = 3,
    EMC_STAGE_COMPRESS = 4,
    EMC_STAGE_TOTAL = 5,
    EMC_STAGE_LOGL_GEMM = 6,
    EMC_STAGE_MSTEP_GEMM = 7,
    EMC_STAGE_PROPOSE = 8, /* proposal: network inference, cell selection, grouped work lists */
    EMC_STAGE_TRAIN = 9,   /* proposal: cell targets, synthetic patterns, network training */
    EMC_N_STAGES = 10
};

/* Search used by an iteration (emc_iter_stats.search). */
enum { EMC_SEARCH_FULL = 0, EMC_SEARCH_PROPOSAL = 1 };

typedef struct emc_iter_stats {
    double total_logl; /* sum_i lse_i, FP64 (the convergence measure) */
    double sigma;      /* annealed width used (main.cpp:251) */
    double fluence;    /* global fluence used (TRUE_POISSON) */
    double model_avg;  /* normaliser: mean of the divided model over the voxels with weight (Wt > 0;
                          the reference averages its positive voxels, emc_cuda.cu:639-661) */
    double weight_sum; /* sum of the insertion weights Wt */
    float stage_ms[EMC_N_STAGES]; /* HIP-event times when options.timing = 1 */
    /* learned proposal (zero / -1 when it is not enabled) */
    int32_t search;        /* EMC_SEARCH_* */
    int32_t n_fallback;    /* patterns searched over every cell (predicted mass < min_mass) */
    double predicted_mass; /* mean over patterns of the network's mass in the kept cells */
    double captured_mass;  /* full search with a trained network: mean over patterns of the
                              responsibility mass inside the cells the network would keep; else -1 */
    double train_loss;     /* mean cross-entropy of the last training pass; -1 if none ran */
    double work_fraction;  /* scored (rotation, pattern) pairs / (S * I) */
    double logits_skipped; /* fraction of the (256-rotation band, pattern) logit blocks whose stores
                              the logL kernel left out (options.keep_logits = 0); 0 otherwise */
} emc_iter_stats;

typedef struct emc_ctx emc_ctx;

/* ---- context ----------------------------------------------------------------------------- */

/* Defaults: TRUE_POISSON, TRILINEAR, BF16X2, BINNED, sigma 0.4 -> 0.12 over 25 (emc.conf:8-10). */
void emc_default_options(emc_options *opt);

/* Replaces the allocation block of main.cpp:158-209 and cuda_set_device (emc_cuda.cu:866-869).
 * Allocates every device buffer for dims.  The iteration and its stages (emc_iterate, emc_logl,
 * emc_normalize_resp, emc_maximize, emc_compress, emc_normalize_model), with or without the
 * proposal, allocate nothing; the input setters (the insertion lists of a new geometry, a count
 * operand a new pattern set needs), emc_expand's staging buffer and emc_proposal_enable do. */
emc_status emc_create(const emc_dims *dims, const emc_options *opt, emc_ctx **out);
void emc_destroy(emc_ctx *ctx);
const char *emc_last_error(const emc_ctx *ctx); /* ctx may be NULL: last create/host error */
emc_status emc_device_info(int32_t device, char *name, int32_t name_len, int32_t *n_cu,
                           int64_t *hbm_bytes); /* cuda_print_device_info (emc_cuda.cu:800-810) */

/* ---- inputs (host -> device) --------------------------------------------------------------- */

/* cuda_allocate_rotations (emc_cuda.cu:695-701) + cuda_copy_weight_to_device (emc_cuda.cu:903-907).
 * q: [S][4] unit quaternions, w: [S] weights (> 0). */
emc_status emc_set_rotations(emc_ctx *ctx, const float *q, const float *w);
/* cuda_allocate_coords (emc_cuda.cu:718-726).  Each [L*L], index y*L + x. */
emc_status emc_set_coords(emc_ctx *ctx, const float *x, const float *y, const float *z);
/* cuda_allocate_mask (emc_cuda.cu:689-693).  [L*L], values 0 or 1. */
emc_status emc_set_mask(emc_ctx *ctx, const int32_t *mask);
/* cuda_allocate_images (emc_cuda.cu:703-709).  [I][L*L] photon counts (>= 0).  Must follow
 * emc_set_mask.  On the MFMA modes integer counts up to 256 are one exact BF16 term; larger ones
 * and non-integer values (binned means, normalised images) are a BF16 hi + lo pair (exact for
 * integers up to 65535, else within 2^-17 relative); the chi^2 kinds' squared counts likewise. */
emc_status emc_set_patterns_i32(emc_ctx *ctx, const int32_t *counts);
emc_status emc_set_patterns_f32(emc_ctx *ctx, const float *counts);
/* cuda_allocate_model / cuda_copy_model_2_device (emc_cuda.cu:675-685).  [L^3]. */
emc_status emc_set_model(emc_ctx *ctx, const float *model);

/* ---- outputs (device -> host) -------------------------------------------------------------- */

/* cuda_copy_model (emc_cuda.cu:608-610), main.cpp:479-480.  Either pointer may be NULL. */
emc_status emc_get_model(emc_ctx *ctx, float *model, float *weight);
/* cuda_model_max / cuda_model_average (emc_cuda.cu:629-671). */
emc_status emc_model_stats(emc_ctx *ctx, double *max, double *average);
/* Logits of the last emc_logl, [S][I], constants included (the reference's d_respons before
 * normalisation, emc_cuda.cu:270-275). */
emc_status emc_get_logl(emc_ctx *ctx, float *logits);
/* Responsibilities [S][I] after emc_normalize_resp (the isDebug dump, main.cpp:395-417).  With
 * options.scaling on an MFMA precision the device keeps P / phi for the M-step; this returns P. */
emc_status emc_get_resp(emc_ctx *ctx, float *resp);
/* The same for the block of rotations [s0, s0+ns) and patterns [i0, i0+ni): out [ns][ni] (a window
 * of the isDebug dump, for inspection at sizes where [S][I] does not fit the host). */
emc_status emc_get_logl_block(emc_ctx *ctx, int32_t s0, int32_t ns, int32_t i0, int32_t ni, float *logits);
emc_status emc_get_resp_block(emc_ctx *ctx, int32_t s0, int32_t ns, int32_t i0, int32_t ni, float *resp);
/* Per-pattern log-sum-exp [I] of the last E-step, full or proposal (main.cpp:384-391, FP64). */
emc_status emc_get_lse(emc_ctx *ctx, double *lse);
/* Slice sums of the last M-step: T [S] (emc_atomic.cu:114-122). */
emc_status emc_get_slice_weights(emc_ctx *ctx, double *T);

/* ---- stages (main.cpp:239-480) ------------------------------------------------------------- */

/* cuda_get_slices (emc_cuda.cu:542-560): slices of rotations [s0, s0+count) into out
 * [count][L*L]; invalid samples (and masked pixels) are -1. */
emc_status emc_expand(emc_ctx *ctx, int32_t s0, int32_t count, float *out);
/* Slicing + likelihood for all rotations: the chunk loop of main.cpp:351-365
 * (cuda_get_slices + cuda_calculate_responsabilities, emc_cuda.cu:280-295). */
emc_status emc_logl(emc_ctx *ctx, float sigma);
/* cuda_max_vector + cuda_respons_max_expf + host log/add + cuda_norm_respons_sumexpf
 * (main.cpp:367-392).  total_logl (may be NULL) receives sum_i lse_i. */
emc_status emc_normalize_resp(emc_ctx *ctx, double *total_logl);
/* Replaces the responsibilities (for stage tests / external E-steps): [S][I].  With
 * options.scaling the M-step then uses P / phi with the phi of the last likelihood pass (1 before
 * any). */
emc_status emc_set_resp(emc_ctx *ctx, const float *resp);
/* Replaces the M-step sums (model numerator and weight, [L^3] each) for stage tests and external
 * M-steps, the state emc_compress starts from (the reference's d_model_updated and d_weight after
 * the insertion loop, main.cpp:443-462). */
emc_status emc_set_mstep_sums(emc_ctx *ctx, const float *numerator, const float *weight);
/* cuda_reset_model x2 + chunk loop of cuda_update_slices (main.cpp:443-462). */
emc_status emc_maximize(emc_ctx *ctx);
/* cuda_divide_model_by_weight + cuda_normalize_model + model blur (main.cpp:477-478).  The model is
 * divided by its mean over the voxels that received weight (Wt > 0), which, unlike the reference's
 * mean over positive voxels, does not depend on where responsibilities underflow; the blur then
 * runs on the normalised model (negative = missing voxels keep their value). */
emc_status emc_compress(emc_ctx *ctx, double *model_avg);
/* cuda_normalize_model (emc_cuda.cu:788-798) on the current model (main.cpp:180): mean over the
 * positive voxels, as the reference (an initial model has no weights). */
emc_status emc_normalize_model(emc_ctx *ctx, double *model_avg);
/* One EMC iteration, main.cpp:239-480: sigma schedule, logl, normalise, maximize, compress.
 * stats may be NULL.  Synchronises before returning. */
emc_status emc_iterate(emc_ctx *ctx, int32_t iteration, emc_iter_stats *stats);

/* ---- allocation-free entry point -------------------------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  emc_compress's divide,
 * normalisation (mean over the voxels with weight) and blur, on caller-owned device buffers of one
 * L^3 model on the current device: model = blur(numerator / weight / avg), -1/avg where weight is 0;
 * *avg (device pointer, may be NULL) receives the normaliser.  stream is a hipStream_t (NULL: the
 * default stream); the call only enqueues.  Allocates nothing: its scratch (two L^3 volumes for
 * L <= 128, about 16 MiB) is a static per-device array that every call on that device shares, so
 * calls that overlap on one device race.  L outside 4..128 is EMC_E_INVALID. */
emc_status emc_compress_device(int32_t L, double blur_sigma, const float *numerator, const float *weight,
                               float *model, double *avg, void *stream);

/* ---- host utilities (no GPU needed) -------------------------------------------------------- */

/* n_to_samples(n)/2 (rotations.c:43, :457). */
int32_t emc_rotation_count(int32_t n);
/* generate_rotation_list (rotations.c:77-458): same order, FP32 rounding, pruning and weights.
 * q [S][4], w [S] with S = emc_rotation_count(n). */
emc_status emc_generate_rotations(int32_t n, float *q, float *w);
/* emc_helper::calculate_coordinates (emc_helper.cpp:55-79). */
emc_status emc_coords_ewald(int32_t L, float pixel_size, float detector_distance,
                            float wavelength, float *x, float *y, float *z);
/* Flat detector plane (Z = 0) of the synthetic workloads. */
emc_status emc_coords_flat(int32_t L, float *x, float *y, float *z);
/* Annular mask rb <= r <= ro on pixel-centre radius; returns the active count (< 0 on error). */
int32_t emc_mask_annulus(int32_t L, double rb, double ro, int32_t *mask);
/* Annealed width of main.cpp:251. */
double emc_sigma_schedule(int32_t iteration, double sigma_start, double sigma_final,
                          int32_t half_life);

/* Synthetic data (no reference counterpart; the reference reads HDF5, emc_helper.cpp:198-248).
 * Object: n_atoms Gaussian atoms (sigma 1 voxel, unit peak) uniform in a sphere of `radius`
 * voxels on an L^3 grid.  Model: fftshifted |DFT|^2 of the object, mean 1. */
emc_status emc_synth_object(int32_t L, int32_t n_atoms, double radius, uint64_t seed,
                            float *object);
emc_status emc_synth_model(int32_t L, int32_t n_atoms, double radius, uint64_t seed, float *model);
/* Patterns: a seeded uniform random quaternion per pattern (Shoemake, as rotations.c:30-41),
 * a trilinear slice of `model` at coords, scaled to `photons` expected counts over mask == 1
 * pixels, Poisson-sampled.  counts [I][L*L] (0 on masked pixels), q_true [I][4] (may be NULL). */
emc_status emc_synth_patterns(const float *model, int32_t L, const float *x, const float *y,
                              const float *z, const int32_t *mask, int32_t n_patterns,
                              double photons, uint64_t seed, int32_t *counts, float *q_true);
/* Uniform random [0,1) model (model_input = 0, emc_helper.cpp:449-456, own RNG). */
emc_status emc_synth_uniform_model(int32_t L, uint64_t seed, float *model);

/* model_input = 3 (emc_helper.cpp:457-500): radial average of the patterns [I][L*L] over integer
 * pixel-radius shells, linearly interpolated onto voxel radii < L/2; -1 outside.  L even. */
emc_status emc_model_init_radial(int32_t L, const float *patterns, int32_t n_patterns, float *model);
/* model_input = 1 (emc_helper.cpp:501-516 with insert_slice, :80-123): each pattern inserted at a
 * seeded uniform random orientation (nearest voxel round(L/2 + 0.5 + R x), mask == 1 pixels, weight
 * 1), then divided by the weight (0 where none).  q_out [I][4] receives the orientations (may be
 * NULL). */
emc_status emc_model_init_orientations(int32_t L, const float *patterns, int32_t n_patterns,
                                       const int32_t *mask, const float *x, const float *y,
                                       const float *z, uint64_t seed, float *q_out, float *model);
/* model_input = 2 (emc_helper.cpp:517-531): |model| from a file, here raw float32 [L^3] in voxel
 * order ((z * L) + y) * L + x (the `emc` driver's output format) instead of libspimage HDF5. */
emc_status emc_model_read_raw(const char *path, int32_t L, float *model);
/* normalize_images (main.cpp:96-97, emc_helper.cpp:389-419): each pattern [P] scaled by
 * P / sum_{mask == 1} of its values (left as it is when that sum is 0).  out may equal patterns. */
emc_status emc_normalize_patterns(const float *patterns, int32_t n_patterns, int32_t P,
                                  const int32_t *mask, float *out);

/* ---- learned orientation proposal (no reference counterpart: main.cpp:239-480 always scores
 * every rotation for every pattern) -------------------------------------------------------------
 *
 * The context's rotations are the fine grid.  A coarse grid (e.g. emc_generate_rotations(2),
 * 420 rotations) partitions it into cells: fine rotation s belongs to the coarse rotation c with
 * the largest |q_s . q_c|.  A network  x -> 512 -> 512 -> C  (GELU, softmax over cells; x = log1p
 * of the pattern pooled onto 32 x 32, standardised) predicts each pattern's responsibility mass
 * per cell.  It is trained on the device (BF16 MFMA forward/backward, FP32 master weights and
 * Adam) on the cell sums of the responsibilities that full-search iterations produce, plus
 * patterns sliced from the current model and Poisson-sampled.  A proposal iteration scores, for
 * each pattern, only the fine rotations of its top_k cells (all cells when the predicted mass in
 * them is below min_mass) as a grouped MFMA contraction; its softmax, M-step sum and insertion
 * see P = 0 elsewhere.  Iterations 0 .. warmup_full-1 and every full_every-th iteration after
 * them search fully, train the network and measure the captured mass; after a proposal
 * iteration's compress the network takes refresh_epochs passes over patterns sliced from the new
 * model.  Needs an MFMA precision, TRUE_POISSON, slice_chunk covering all rotations and counts
 * <= 256. */
typedef struct emc_proposal_options {
    int32_t top_k;          /* cells kept per pattern (8) */
    int32_t warmup_full;    /* leading full-search iterations (3) */
    int32_t full_every;     /* after warm-up, every full_every-th iteration searches fully (5) */
    int32_t hidden;         /* hidden width, multiple of 64 (512) */
    int32_t epochs;         /* passes over the training set after each full search (4) */
    int32_t batch;          /* minibatch rows, multiple of 64 (1024) */
    int32_t synth_patterns; /* model-sliced training patterns per pass; -1 = n_patterns / 4 */
    int32_t refresh_epochs; /* passes over fresh model-sliced patterns after each proposal
                               iteration's compress, so the network follows the model (1) */
    int32_t recall;         /* cells searched beside the network's top_k: the pattern's cells of
                               largest max logit in its most recent search (all cells after a
                               full search, the searched ones after a proposal iteration), so
                               a runner-up orientation that overtakes the best is still found;
                               0..8 (8) */
    int32_t train_patterns; /* 1: training after a full search uses the patterns' cell targets
                               besides the model-sliced patterns; 0: model-sliced patterns only,
                               so the network learns the current model's orientation map rather
                               than the previous assignments (1) */
    float min_mass;        /* per-pattern predicted mass below which it is searched fully (0.99) */
    float lr;               /* Adam step size (1.5e-3; beta 0.9 / 0.999, eps 1e-8) */
    float max_fallback;     /* fraction of fully searched patterns above which the iteration
                               runs a full search instead (0.1) */
    float refresh_fraction; /* share of synth_patterns drawn for each refresh pass (0.25) */
    float train_below;      /* after the warm-up, a full search trains the network only when the
                               responsibility mass its prediction captured (measured by that
                               full search) is below this; 1 trains at every full search, 0 never
                               (0.9999) */
    uint64_t seed;          /* initial weights, shuffling, synthetic patterns */
} emc_proposal_options;

void emc_proposal_defaults(emc_proposal_options *o);
/* Coarse cell of each fine rotation: argmax_c |fine_s . coarse_c| (lowest c on ties).  Host. */
emc_status emc_coarse_partition(const float *fine, int32_t n_fine, const float *coarse,
                                int32_t n_coarse, int32_t *cell);
/* Enable the proposal on a context whose rotations, coords, mask and patterns are set; coarse
 * [n_coarse][4].  Computes the pattern features, initialises the network (He-uniform) and Adam
 * state and restarts the iteration counter. */
emc_status emc_proposal_enable(emc_ctx *ctx, const float *coarse, int32_t n_coarse,
                               const emc_proposal_options *opt);
emc_status emc_proposal_disable(emc_ctx *ctx);
/* Network prediction for every pattern: cells [I][top_k] (ascending), mass [I] = the predicted
 * probability of every cell a proposal iteration would search (the top_k and the recalled
 * ones).  Either may be NULL. */
emc_status emc_proposal_predict(emc_ctx *ctx, int32_t *cells, float *mass);
/* Per-pattern captured mass of the last full-search iteration that measured it (the
 * emc_iter_stats.captured_mass average; the cells counted are the top_k predicted ones and the
 * recalled ones): cap [I]. */
emc_status emc_proposal_captured(emc_ctx *ctx, double *cap);
/* Recalled cells of every pattern, most massive first, -1 for none: cells [I][recall]. */
emc_status emc_proposal_recall(emc_ctx *ctx, int32_t *cells);
/* Cell sums of the current full responsibilities: targets [I][n_coarse]. */
emc_status emc_proposal_cell_targets(emc_ctx *ctx, float *targets);
/* Train on the current full responsibilities (cell targets) plus synthetic patterns for
 * `epochs` passes; mean cross-entropy of the last pass in *loss. */
emc_status emc_proposal_train(emc_ctx *ctx, int32_t epochs, double *loss);
/* Pooled log1p features of every pattern before standardisation: raw [I][1024]. */
emc_status emc_proposal_features(emc_ctx *ctx, float *raw);
/* Network parameters (FP32 masters), layouts W1 [H][1024], b1 [H], W2 [H][H], b2 [H],
 * W3 [C][H], b3 [C].  set also refreshes the BF16 copies and clears the Adam state. */
emc_status emc_mlp_get_params(emc_ctx *ctx, float *W1, float *b1, float *W2, float *b2,
                              float *W3, float *b3);
emc_status emc_mlp_set_params(emc_ctx *ctx, const float *W1, const float *b1, const float *W2,
                              const float *b2, const float *W3, const float *b3);
/* One forward pass on given (already standardised) features x [B][1024] -> probs [B][C]. */
emc_status emc_mlp_forward(emc_ctx *ctx, const float *x, int32_t B, float *probs);
/* Gradients of the mean cross-entropy against targets t [B][C] (same layouts as the
 * parameters); *loss = the loss.  Parameters unchanged. */
emc_status emc_mlp_gradients(emc_ctx *ctx, const float *x, const float *t, int32_t B, double *loss,
                             float *gW1, float *gb1, float *gW2, float *gb2, float *gW3,
                             float *gb3);
/* One Adam step with the gradients of (x, t) (step counter advances). */
emc_status emc_mlp_train_step(emc_ctx *ctx, const float *x, const float *t, int32_t B,
                              double *loss);

/* ---- configuration (emc.conf) ----------------------------------------------------------------- */

/* Every key of the reference's Configuration (configs.h:12-86, read at configs.cpp:51-88) with
 * its meaning, plus this engine's keys.  Strings are NUL-terminated. */
typedef struct emc_config {
    int32_t model_side, read_stride, detector_size, rotations_n, sigma_half_life, slice_chunk;
    int32_t N_images, max_iterations;
    double wavelength, pixel_size, detector_distance, sigma_start, sigma_final;
    int32_t diff; /* from diff_type: "absolute", "poisson" (chi^2), "relative", "true_poisson" */
    int32_t normalize_images, known_intensity, model_input, exclude_images;
    int32_t output_period, output_loop, isOutput, isDebug, blur_image, nGPU;
    double blur_sigma, exclude_ratio, model_blur;
    char image_prefix[512], mask_file[512], model_file[512];
    /* engine keys: precision "fp32"|"bf16x2"|"bf16x3", interp "nearest"|"trilinear",
     * insert "atomic"|"fused"|"binned", fluence (0 = automatic), detector "flat"|"ewald" */
    int32_t precision, interp, insert, flat_detector;
    double fluence;
    /* synthetic data (image_prefix = "synthetic"): atoms, sphere radius, photons per pattern,
     * seed, mask beamstop and outer radius (pixels) */
    int32_t synth_atoms;
    double synth_radius, synth_photons, beamstop, outer_radius;
    uint64_t synth_seed;
    /* learned proposal: coarse grid sampling n (0 = off; 2 gives 420 cells) and cells kept per
     * pattern (emc_proposal_options.top_k); the other proposal options keep their defaults */
    int32_t proposal_coarse_n, proposal_top_k;
} emc_config;

/* Defaults for every key (the reference leaves missing keys uninitialised). */
emc_status emc_config_defaults(emc_config *cfg);
/* Config::read_configuration_file (configs.cpp:40-90): libconfig syntax subset; an unknown key,
 * a wrong value type or a bad enum string is EMC_E_INVALID with the line in emc_last_error(NULL). */
emc_status emc_config_parse(const char *text, emc_config *cfg);
emc_status emc_config_read(const char *path, emc_config *cfg);

#ifdef __cplusplus
}
#endif

#endif /* EMC_H */
