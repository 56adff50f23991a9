// Learned orientation proposal: the network that maps a pattern to a distribution over coarse
// rotation cells, its training on the device (forward, backward, Adam) and the kernels that
// feed it (features, cell targets, synthetic patterns, top-k cells).
//
// No reference counterpart (main.cpp:239-480 scores every rotation for every pattern).
//
// Network: x = standardised log1p of the pattern pooled onto 32 x 32 (D = 1024) -> H -> H -> C,
// GELU (erf form) after both hidden layers, softmax over the C coarse cells, mean cross-entropy
// against the cells' summed responsibilities.  BF16 weights and activations feed MFMA
// (v_mfma_f32_32x32x16_bf16) with FP32 accumulation; FP32 master weights and Adam moments.
//
// GEMM kernel: C[m, n] = sum_k A(m, k) B(n, k), 64 x 64 tile, 256 threads = 4 wavefronts of one
// 32 x 32 MFMA tile each, BK = 128 (64 where K is not a multiple of 128).  Each operand is either
// k-contiguous ("row": X[m][k]) or m-contiguous ("col": X[k][m], a transposed read, used where the
// reduction runs over the batch or over a weight's output index); col operands are transposed
// while staged into LDS.  The network's products are small (a minibatch of 1,024 rows, K of 448 to
// 1,024: 56-128 workgroups), so a launch is bound by the latency of its global loads: the next
// k-step's tile is loaded into registers while the current one is multiplied, and the wide k-step
// keeps the number of such round trips at K / BK.
#include <algorithm>
#include <cmath>

#include "../emc_internal.h"
#include "geom.h"

namespace emc {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int GT = 64, GK = 64;  // tile; K must be a multiple of GK (D, H, Cp and the batch are)
constexpr int kD = 1024;

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float x)
{
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * expf(-0.5f * x * x) * 0.3989422804014327f;
}

struct MlpEpi {
    float *out32;       // FP32 output [M][ldo] (pre-activation / logits / gradient); may be null
    bf16_t *out16;      // BF16 output [M][ldo] (activation / back-propagated delta)
    int ldo;
    const float *bias;  // [N]
    const float *pre;   // [M][ldo] pre-activation for GELU'
    int m_valid, n_valid;
};

// A(m,k) at A[m*lda + k] (row) or A[k*lda + m] (col); likewise B(n,k).  K % BK == 0.
template <bool ACOL, bool BCOL, int EPI, int BK>
__global__ __launch_bounds__(256) void k_mlp_gemm(const bf16_t *__restrict__ A, int lda, const bf16_t *__restrict__ B,
                                                  int ldb, int K, MlpEpi e)
{
    static_assert(BK == 64 || BK == 128, "a 64 x BK tile is BK / 32 chunks of 8 elements per thread");
    constexpr int LD = BK + 8;  // padded row: 16-byte fragment reads of 32 consecutive rows hit distinct banks
    constexpr int NC = BK / 32;  // chunks of 8 elements per thread and operand
    __shared__ __attribute__((aligned(16))) bf16_t As[GT][LD], Bs[GT][LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    // chunk q of this thread: row operand -> row id / (BK / 8), k offset (id % (BK / 8)) * 8 with
    // id = q * 256 + tid; col operand -> k row NC (tid / 8) + q, 8 consecutive m (or n) from
    // (tid % 8) * 8, so that the thread holds NC consecutive k of each of its 8 m
    auto load = [&](const bf16_t *X, int ld, int x0, bool col, int k0, u32x4 *r) {
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const int id = q * 256 + tid;
            if (!col)
                r[q] = *reinterpret_cast<const u32x4 *>(X + (size_t)(x0 + id / (BK / 8)) * ld + k0 + (id % (BK / 8)) * 8);
            else
                r[q] = *reinterpret_cast<const u32x4 *>(X + (size_t)(k0 + NC * (tid >> 3) + q) * ld + x0 + (tid & 7) * 8);
        }
    };
    auto stage = [&](bf16_t (*Xs)[LD], bool col, const u32x4 *r) {
        if (!col) {
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                const int id = q * 256 + tid;
                *reinterpret_cast<u32x4 *>(&Xs[id / (BK / 8)][(id % (BK / 8)) * 8]) = r[q];
            }
        } else {  // transposed into LDS: the NC k of one m as one 2 NC-byte write
            const int kq = NC * (tid >> 3), mm = (tid & 7) * 8;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                uint32_t h[NC];  // element j of each chunk (k = kq + q)
#pragma unroll
                for (int q = 0; q < NC; ++q) h[q] = (j & 1) ? (r[q][j >> 1] >> 16) : (r[q][j >> 1] & 0xffffu);
                if constexpr (NC == 4)
                    *reinterpret_cast<uint2 *>(&Xs[mm + j][kq]) = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
                else
                    *reinterpret_cast<uint32_t *>(&Xs[mm + j][kq]) = h[0] | (h[1] << 16);
            }
        }
    };
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    u32x4 ra[NC], rb[NC];
    load(A, lda, m0, ACOL, 0, ra);
    load(B, ldb, n0, BCOL, 0, rb);
    const int fr = lane & 31, fk = (lane >> 5) * 8;
    for (int k0 = 0; k0 < K; k0 += BK) {
        stage(As, ACOL, ra);
        stage(Bs, BCOL, rb);
        __syncthreads();
        if (k0 + BK < K) {  // the next k-step's loads are in flight during this step's products
            load(A, lda, m0, ACOL, k0 + BK, ra);
            load(B, ldb, n0, BCOL, k0 + BK, rb);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[wm + fr][kk + fk]);
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&Bs[wn + fr][kk + fk]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // accumulator: lane holds column n = lane % 32 and rows 8*(r/4) + 4*(lane/32) + r%4
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int n = n0 + wn + (lane & 31);
        if (m >= e.m_valid || n >= e.n_valid) continue;
        const size_t o = (size_t)m * e.ldo + n;
        float v = acc[r];
        if (EPI == MLP_EPI_STORE) {
            e.out32[o] = v;
        } else if (EPI == MLP_EPI_BIAS_GELU) {
            v += e.bias[n];
            if (e.out32) e.out32[o] = v;
            e.out16[o] = f2bf(gelu(v));
        } else if (EPI == MLP_EPI_BIAS) {
            e.out32[o] = v + e.bias[n];
        } else {  // delta of the layer below: (dY . W) * GELU'(pre)
            e.out16[o] = f2bf(v * gelu_d(e.pre[o]));
        }
    }
}

// Pooled features: counts of rows [0, rows) summed over f x f pixel blocks onto 32 x 32, log1p;
// rows in [rows, rows_pad) are zero.  K8 / K16 / K32 (one non-null) are [row][ld] over the
// compacted active pixels, pix maps them to the detector.
__global__ __launch_bounds__(256) void k_features(const float *__restrict__ K32, const bf16_t *__restrict__ K16,
                                                  const uint8_t *__restrict__ K8, int ld, int Pa,
                                                  const int32_t *__restrict__ pix, int L, int f, int rows,
                                                  float *__restrict__ raw)
{
    __shared__ float acc[kD];
    const int b = blockIdx.x;
    for (int k = threadIdx.x; k < kD; k += 256) acc[k] = 0.0f;
    __syncthreads();
    if (b < rows) {
        for (int j = threadIdx.x; j < Pa; j += 256) {
            const size_t o = (size_t)b * ld + j;
            const float k = K32 ? K32[o] : (K8 ? (float)K8[o] : bf2f(K16[o]));
            const int p = pix[j], x = (p % L) / f, y = (p / L) / f;
            if (k != 0.0f && x < 32 && y < 32) atomicAdd(&acc[y * 32 + x], k);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kD; k += 256) raw[(size_t)b * kD + k] = b < rows ? log1pf(acc[k]) : 0.0f;
}

// Per-feature mean and 1 / standard deviation over rows (FP64 sums; fixed order).
__global__ __launch_bounds__(256) void k_feat_stats(const float *__restrict__ raw, int rows, float *__restrict__ mu,
                                                    float *__restrict__ istd)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= kD) return;
    double s = 0.0, s2 = 0.0;
    for (int b = 0; b < rows; ++b) {
        const double v = raw[(size_t)b * kD + k];
        s += v;
        s2 += v * v;
    }
    const double m = s / rows, var = fmax(s2 / rows - m * m, 0.0);
    mu[k] = (float)m;
    istd[k] = (float)(1.0 / sqrt(var + 1e-4));
}

__global__ void k_standardise(const float *__restrict__ raw, int64_t n_valid, int64_t n, const float *__restrict__ mu,
                              const float *__restrict__ istd, bf16_t *__restrict__ X)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(t % kD);
        X[t] = f2bf(t < n_valid ? (raw[t] - mu[k]) * istd[k] : 0.0f);
    }
}

// Row-wise softmax over the first C of Cp logits -> p (FP32).  With targets t: per-row
// cross-entropy -sum t log p (FP64) and dz = (p - t) * scale (BF16); rows >= rows_valid give 0.
__global__ __launch_bounds__(64) void k_softmax_ce(const float *__restrict__ z, int Cp, int C, int rows_valid,
                                                   const float *__restrict__ t, float scale, float *__restrict__ p,
                                                   bf16_t *__restrict__ dz, double *__restrict__ loss_part)
{
    const int b = blockIdx.x, lane = threadIdx.x;
    float m = -3.0e38f;
    for (int c = lane; c < C; c += 64) m = fmaxf(m, z[(size_t)b * Cp + c]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    float s = 0.0f;
    for (int c = lane; c < C; c += 64) s += expf(z[(size_t)b * Cp + c] - m);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    const double ls_row = log((double)s);
    double ls = 0.0;
    const bool valid = b < rows_valid;
    for (int c = lane; c < Cp; c += 64) {
        const size_t o = (size_t)b * Cp + c;
        const float pv = c < C ? expf(z[o] - m) / s : 0.0f;
        p[o] = pv;
        if (t) {
            const float tv = c < C ? t[o] : 0.0f;
            if (valid && tv > 0.0f) ls -= (double)tv * ((double)(z[o] - m) - ls_row);
            dz[o] = f2bf(valid ? (pv - tv) * scale : 0.0f);
        }
    }
    if (t) {
        for (int o = 32; o > 0; o >>= 1) ls += __shfl_xor(ls, o, 64);
        if (lane == 0) loss_part[b] = ls;
    }
}

// acc += sum of x[0..n) (one workgroup, fixed order).
__global__ __launch_bounds__(256) void k_accum_f64(const double *__restrict__ x, int n, double *__restrict__ acc)
{
    __shared__ double red[4];
    double a = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) a += x[k];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) *acc += red[0] + red[1] + red[2] + red[3];
}

// Column sums of a BF16 matrix [rows][ld] -> g[n] (bias gradients).  A block covers 32 columns
// with 8 row groups (rows g, g + 8, ...; eight loads in flight per thread), combined in LDS in a
// fixed order.
__global__ __launch_bounds__(256) void k_colsum(const bf16_t *__restrict__ d, int rows, int ld, int n_valid,
                                                float *__restrict__ g)
{
    __shared__ float part[8][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5, n = blockIdx.x * 32 + tx;
    float acc = 0.0f;
    if (n < n_valid) {
        int b = ty;
        for (; b + 56 < rows; b += 64) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = bf2f(d[(size_t)(b + 8 * u) * ld + n]);
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += v[u];
        }
        for (; b < rows; b += 8) acc += bf2f(d[(size_t)b * ld + n]);
    }
    part[ty][tx] = acc;
    __syncthreads();
    if (ty == 0 && n < n_valid) {
        float s = 0.0f;
#pragma unroll
        for (int r = 0; r < 8; ++r) s += part[r][tx];
        g[n] = s;
    }
}

// Adam (Kingma & Ba 2015) on FP32 masters, BF16 copy refreshed.  c1, c2 = bias corrections.
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, bf16_t *__restrict__ w16, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const float gk = g[k];
        const float mk = b1 * m[k] + (1.0f - b1) * gk;
        const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
        m[k] = mk;
        v[k] = vk;
        const float wk = w[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
        w[k] = wk;
        w16[k] = f2bf(wk);
    }
}

__global__ void k_to_bf16(const float *__restrict__ w, int64_t n, bf16_t *__restrict__ w16)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        w16[k] = f2bf(w[k]);
}

// Top-k cells per pattern (ascending cell order; larger probability first, lower cell on ties)
// and the predicted mass they hold.  k <= 64.
__global__ __launch_bounds__(64) void k_topk(const float *__restrict__ p, int Cp, int C, int k,
                                             int32_t *__restrict__ cells, float *__restrict__ mass)
{
    const int b = blockIdx.x, lane = threadIdx.x;
    __shared__ int chosen[64];
    float tot = 0.0f;
    for (int r = 0; r < k; ++r) {
        float best = -1.0f;
        int bi = 0x7fffffff;
        for (int c = lane; c < C; c += 64) {
            bool used = false;
            for (int q = 0; q < r; ++q) used |= chosen[q] == c;
            const float v = p[(size_t)b * Cp + c];
            if (!used && (v > best || (v == best && c < bi))) {
                best = v;
                bi = c;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const float ov = __shfl_xor(best, o, 64);
            const int oi = __shfl_xor(bi, o, 64);
            if (ov > best || (ov == best && oi < bi)) {
                best = ov;
                bi = oi;
            }
        }
        if (lane == 0) chosen[r] = bi;
        __syncthreads();
        tot += best;
    }
    if (lane == 0) {
        for (int a = 1; a < k; ++a) {  // insertion sort of the k chosen cells
            const int v = chosen[a];
            int q = a - 1;
            while (q >= 0 && chosen[q] > v) {
                chosen[q + 1] = chosen[q];
                --q;
            }
            chosen[q + 1] = v;
        }
        for (int a = 0; a < k; ++a) cells[(size_t)b * k + a] = chosen[a];
        mass[b] = tot;
    }
}

// Predicted mass of the cells a proposal iteration searches for each pattern: the top-k cells'
// (already in mass[i]) plus the recalled cells' not among them.
__global__ void k_union_mass(const float *__restrict__ p, int Cp, int I, const int32_t *__restrict__ cells, int k,
                             const int32_t *__restrict__ recall, int R, float *__restrict__ mass)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= I) return;
    const int32_t *ci = cells + (size_t)i * k, *ri = recall + (size_t)i * R;
    float a = mass[i];
    for (int q = 0; q < R; ++q) {
        const int cr = ri[q];
        bool seen = cr < 0;
        for (int u = 0; u < k; ++u) seen |= ci[u] == cr;
        for (int u = 0; u < q; ++u) seen |= ri[u] == cr;
        if (!seen) a += p[(size_t)i * Cp + cr];
    }
    mass[i] = a;
}

// Cell sums of the full responsibilities: Tt[c][i] = sum over rotations s of cell c of
// (Ph + Pl)[s][i]; with LOGITS also the cell's largest logit Mx[c][i] = max_s Lg[s][i] (-inf for
// an empty cell).  Grid (pattern blocks, cells); rows are read coalesced.  pdirty (null: read
// every row) marks the 256-pattern row segments of P that may hold a non-zero; the others hold
// zeros and are not read (adding their zeros would leave the sums unchanged).
template <bool LOGITS>
__global__ __launch_bounds__(256) void k_cell_targets(const bf16_t *__restrict__ Ph, const bf16_t *__restrict__ Pl,
                                                      const float *__restrict__ Lg, int64_t ld, int I,
                                                      const int32_t *__restrict__ cell_rot,
                                                      const int32_t *__restrict__ rbeg, const uint8_t *__restrict__ pdirty,
                                                      int nseg, float *__restrict__ Tt, float *__restrict__ Mx)
{
    // eight patterns per thread (one segment): 16-byte loads of hi and lo, two of the logits
    // (i0 < I implies i0 + 8 <= ld); each pattern's sum runs over the cell's rotations in order
    const int i0 = (blockIdx.x * 256 + threadIdx.x) * 8, c = blockIdx.y;
    if (i0 >= I) return;
    float acc[8], mx[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        acc[k] = 0.0f;
        mx[k] = -INFINITY;
    }
    for (int g = rbeg[c]; g < rbeg[c + 1]; ++g) {
        const int s = cell_rot[g];
        const size_t o = (size_t)s * ld + i0;
        if (!pdirty || pdirty[(size_t)s * nseg + (i0 >> 8)]) {
            const u32x4 h = *reinterpret_cast<const u32x4 *>(Ph + o), l = *reinterpret_cast<const u32x4 *>(Pl + o);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                acc[k] += bf2f((bf16_t)(h[k >> 1] >> (16 * (k & 1)))) + bf2f((bf16_t)(l[k >> 1] >> (16 * (k & 1))));
        }
        if (LOGITS) {
            const float4 a = *reinterpret_cast<const float4 *>(Lg + o), b = *reinterpret_cast<const float4 *>(Lg + o + 4);
            const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
            for (int k = 0; k < 8; ++k) mx[k] = fmaxf(mx[k], v[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (i0 + k < I) {
            Tt[(size_t)c * ld + i0 + k] = acc[k];
            if (LOGITS) Mx[(size_t)c * ld + i0 + k] = mx[k];
        }
}

// Responsibility mass inside the cells a proposal iteration would search for each pattern: the
// predicted top k and the R recalled ones (-1 = none), each cell counted once; all cells when the
// predicted mass is below min_mass.
__global__ void k_captured(const float *__restrict__ Tt, int64_t ld, int I, int C, const int32_t *__restrict__ cells,
                           const float *__restrict__ mass, int k, float min_mass, const int32_t *__restrict__ recall,
                           int R, double *__restrict__ cap)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= I) return;
    double a = 0.0;
    if (mass[i] < min_mass) {
        for (int c = 0; c < C; ++c) a += Tt[(size_t)c * ld + i];
    } else {
        const int32_t *ci = cells + (size_t)i * k;
        for (int q = 0; q < k; ++q) a += Tt[(size_t)ci[q] * ld + i];
        const int32_t *ri = recall + (size_t)i * R;
        for (int q = 0; q < R; ++q) {
            const int cr = ri[q];
            bool seen = cr < 0;
            for (int u = 0; u < k; ++u) seen |= ci[u] == cr;
            for (int u = 0; u < q; ++u) seen |= ri[u] == cr;
            if (!seen) a += Tt[(size_t)cr * ld + i];
        }
    }
    cap[i] = a;
}

// Minibatch gather: row b takes training item id = idx[b]: a pattern (id < I: features X[id],
// targets Tt[:, id]) or a synthetic pattern (features Xs[id - I], one-hot target on the cell of
// its rotation).
__global__ __launch_bounds__(256) void k_gather_batch(const bf16_t *__restrict__ X, const bf16_t *__restrict__ Xs,
                                                      int I, const float *__restrict__ Tt, int64_t ldt,
                                                      const int32_t *__restrict__ syn_rot,
                                                      const int32_t *__restrict__ cell_of, int C, int Cp,
                                                      const int32_t *__restrict__ idx, bf16_t *__restrict__ Xb,
                                                      float *__restrict__ tb)
{
    const int b = blockIdx.x, id = idx[b];
    const bf16_t *src = id < I ? X + (size_t)id * kD : Xs + (size_t)(id - I) * kD;
    const int cell = id < I ? -1 : cell_of[syn_rot[id - I]];
    for (int k = threadIdx.x; k < kD / 8; k += 256)
        reinterpret_cast<u32x4 *>(Xb + (size_t)b * kD)[k] = reinterpret_cast<const u32x4 *>(src)[k];
    for (int c = threadIdx.x; c < Cp; c += 256) {
        float t = 0.0f;
        if (c < C) t = id < I ? Tt[(size_t)c * ldt + id] : (c == cell ? 1.0f : 0.0f);
        tb[(size_t)b * Cp + c] = t;
    }
}

// Counter-based generator (splitmix64 steps from a hash of seed, pattern and pixel).
struct DevRng {
    uint64_t s;
    __device__ uint64_t next()
    {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    __device__ float uniform() { return (float)(next() >> 40) * 0x1.0p-24f; }  // [0, 1)
};

// Poisson deviate: lam < 10 by multiplying uniforms (Knuth, TAOCP vol. 2, 3.4.1); lam >= 10 by
// PTRS transformed rejection (Hormann 1993).  Both loops are capped (the caps are reached with
// negligible probability); the capped result is the rounded mean.
__device__ float poisson_dev(float lam, DevRng &r)
{
    if (!(lam > 0.0f)) return 0.0f;
    if (lam < 10.0f) {
        const float lim = expf(-lam);
        float prod = r.uniform();
        int k = 0;
        while (prod > lim && k < 200) {
            prod *= r.uniform();
            ++k;
        }
        return (float)k;
    }
    const float slam = sqrtf(lam), loglam = logf(lam);
    const float b = 0.931f + 2.53f * slam;
    const float a = -0.059f + 0.02483f * b;
    const float inv_alpha = 1.1239f + 1.1328f / (b - 3.4f);
    const float vr = 0.9277f - 3.6224f / (b - 2.0f);
    for (int it = 0; it < 64; ++it) {
        const float U = r.uniform() - 0.5f;
        const float V = r.uniform();
        const float us = 0.5f - fabsf(U);
        const float kf = floorf((2.0f * a / us + b) * U + lam + 0.43f);
        if (us >= 0.07f && V <= vr) return kf;
        if (kf < 0.0f || (us < 0.013f && V > us)) continue;
        if (logf(V) + logf(inv_alpha) - logf(a / (us * us) + b) <= -lam + kf * loglam - lgammaf(kf + 1.0f)) return kf;
    }
    return rintf(lam);
}

// Synthetic training patterns: slice of the current model at rotation syn_rot[b], scaled so its
// expected total over the active pixels is `photons`, Poisson-sampled.  Pass 1 writes the slice
// values and sums them (FP64); pass 2 samples in place.
template <int INTERP>
__global__ __launch_bounds__(256) void k_synth_counts(const float *__restrict__ model, int L,
                                                      const float *__restrict__ rotm, const int32_t *__restrict__ syn_rot,
                                                      const float *__restrict__ axyz, int Pa, int Pp, float photons,
                                                      uint64_t seed, float *__restrict__ counts)
{
    __shared__ double red[4];
    const int b = blockIdx.x, s = syn_rot[b];
    float m[9];
    load_rot(rotm, s, m);
    float *row = counts + (size_t)b * Pp;
    double sw = 0.0;
    for (int j = threadIdx.x; j < Pp; j += 256) {
        float W = 0.0f;
        if (j < Pa) W = fmaxf(sample<INTERP>(model, L, m, axyz[j], axyz[Pp + j], axyz[2 * Pp + j]), 0.0f);
        row[j] = W;
        sw += (double)W;
    }
    for (int o = 32; o > 0; o >>= 1) sw += __shfl_xor(sw, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sw;
    __syncthreads();
    const double tot = red[0] + red[1] + red[2] + red[3];
    const float scale = tot > 0.0 ? (float)(photons / tot) : 0.0f;
    for (int j = threadIdx.x; j < Pp; j += 256) {
        const float W = row[j];
        float k = 0.0f;
        if (W > 0.0f) {
            DevRng r{seed ^ ((uint64_t)b * 0xD1B54A32D192ED03ull) ^ ((uint64_t)j * 0xABC98388FB8FAC03ull)};
            k = poisson_dev(W * scale, r);
        }
        row[j] = k;
    }
}

}  // namespace

// ---- host-side launchers ------------------------------------------------------------------

hipError_t mlp_gemm(bool acol, bool bcol, int epi, const bf16_t *A, int lda, const bf16_t *B, int ldb, int M, int N,
                    int K, float *out32, bf16_t *out16, int ldo, const float *bias, const float *pre, int m_valid,
                    int n_valid, hipStream_t st)
{
    if (M % GT || N % GT || K % GK) return hipErrorInvalidValue;
    MlpEpi e{out32, out16, ldo, bias, pre, m_valid, n_valid};
    const dim3 grid(N / GT, M / GT), blk(256);
#define EMC_MLP_CASE(AC, BC, EP)                                                                                   \
    if (acol == AC && bcol == BC && epi == EP) {                                                                   \
        if (K % 128 == 0)                                                                                          \
            hipLaunchKernelGGL((k_mlp_gemm<AC, BC, EP, 128>), grid, blk, 0, st, A, lda, B, ldb, K, e);             \
        else                                                                                                       \
            hipLaunchKernelGGL((k_mlp_gemm<AC, BC, EP, 64>), grid, blk, 0, st, A, lda, B, ldb, K, e);              \
        return hipGetLastError();                                                                                  \
    }
    EMC_MLP_CASE(false, false, MLP_EPI_BIAS_GELU)
    EMC_MLP_CASE(false, false, MLP_EPI_BIAS)
    EMC_MLP_CASE(false, true, MLP_EPI_GELU_GRAD)
    EMC_MLP_CASE(true, true, MLP_EPI_STORE)
#undef EMC_MLP_CASE
    return hipErrorInvalidValue;
}

hipError_t mlp_features(const float *K32, const bf16_t *K16, const uint8_t *K8, int ld, int Pa, const int32_t *pix,
                        int L, int rows, int rows_pad, float *raw, hipStream_t st)
{
    const int f = std::max(1, L / 32);
    hipLaunchKernelGGL(k_features, dim3(rows_pad), dim3(256), 0, st, K32, K16, K8, ld, Pa, pix, L, f, rows, raw);
    return hipGetLastError();
}

hipError_t mlp_feat_stats(const float *raw, int rows, float *mu, float *istd, hipStream_t st)
{
    hipLaunchKernelGGL(k_feat_stats, dim3(kD / 256), dim3(256), 0, st, raw, rows, mu, istd);
    return hipGetLastError();
}

hipError_t mlp_standardise(const float *raw, int rows, int rows_pad, const float *mu, const float *istd, bf16_t *X,
                           hipStream_t st)
{
    const int64_t n = (int64_t)rows_pad * kD;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_standardise, dim3(blocks), dim3(256), 0, st, raw, (int64_t)rows * kD, n, mu, istd, X);
    return hipGetLastError();
}

hipError_t mlp_softmax_ce(const float *z, int Cp, int C, int rows_valid, int rows, const float *t, float scale,
                          float *p, bf16_t *dz, double *loss_part, hipStream_t st)
{
    hipLaunchKernelGGL(k_softmax_ce, dim3(rows), dim3(64), 0, st, z, Cp, C, rows_valid, t, scale, p, dz, loss_part);
    return hipGetLastError();
}

hipError_t mlp_accum(const double *x, int n, double *acc, hipStream_t st)
{
    hipLaunchKernelGGL(k_accum_f64, dim3(1), dim3(256), 0, st, x, n, acc);
    return hipGetLastError();
}

hipError_t mlp_colsum(const bf16_t *d, int rows, int ld, int n_valid, float *g, hipStream_t st)
{
    hipLaunchKernelGGL(k_colsum, dim3((n_valid + 31) / 32), dim3(256), 0, st, d, rows, ld, n_valid, g);
    return hipGetLastError();
}

hipError_t mlp_adam(float *w, const float *g, float *m, float *v, bf16_t *w16, int64_t n, float lr, int t,
                    hipStream_t st)
{
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = (float)(1.0 - std::pow(b1, t)), c2 = (float)(1.0 - std::pow(b2, t));
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_adam, dim3(blocks), dim3(256), 0, st, w, g, m, v, w16, n, lr, (float)b1, (float)b2, 1e-8f,
                       c1, c2);
    return hipGetLastError();
}

hipError_t mlp_to_bf16(const float *w, int64_t n, bf16_t *w16, hipStream_t st)
{
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks), dim3(256), 0, st, w, n, w16);
    return hipGetLastError();
}

hipError_t mlp_topk(const float *p, int Cp, int C, int rows, int k, int32_t *cells, float *mass, hipStream_t st)
{
    if (k < 1 || k > 64 || k > C) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_topk, dim3(rows), dim3(64), 0, st, p, Cp, C, k, cells, mass);
    return hipGetLastError();
}

hipError_t mlp_union_mass(const emc_ctx *c, hipStream_t st)
{
    const Proposal *pr = c->prop;
    if (pr->o.recall == 0) return hipSuccess;
    hipLaunchKernelGGL(k_union_mass, dim3((c->I + 255) / 256), dim3(256), 0, st, pr->prob, pr->Cp, c->I, pr->topc,
                       pr->o.top_k, pr->recall, pr->o.recall, pr->pmass);
    return hipGetLastError();
}

hipError_t mlp_cell_targets(const emc_ctx *c, float *Tt, float *Mx, hipStream_t st)
{
    const Proposal *pr = c->prop;
    const dim3 grid((c->I + 2047) / 2048, pr->C);  // eight patterns per thread
    if (Mx)
        hipLaunchKernelGGL(k_cell_targets<true>, grid, dim3(256), 0, st, c->Ph, c->Pl, c->R, (int64_t)c->Ip, c->I,
                           pr->cell_rot, pr->rbeg, c->pdirty, p_segments(c), Tt, Mx);
    else
        hipLaunchKernelGGL(k_cell_targets<false>, grid, dim3(256), 0, st, c->Ph, c->Pl, nullptr, (int64_t)c->Ip, c->I,
                           pr->cell_rot, pr->rbeg, c->pdirty, p_segments(c), Tt, nullptr);
    return hipGetLastError();
}

hipError_t mlp_captured(const emc_ctx *c, double *cap, hipStream_t st)
{
    const Proposal *pr = c->prop;
    hipLaunchKernelGGL(k_captured, dim3((c->I + 255) / 256), dim3(256), 0, st, pr->Tt, (int64_t)c->Ip, c->I, pr->C,
                       pr->topc, pr->pmass, pr->o.top_k, pr->o.min_mass, pr->recall, pr->o.recall, cap);
    return hipGetLastError();
}

hipError_t mlp_gather_batch(const emc_ctx *c, const int32_t *idx, int rows, hipStream_t st)
{
    const Proposal *pr = c->prop;
    hipLaunchKernelGGL(k_gather_batch, dim3(rows), dim3(256), 0, st, pr->X, pr->Xs, c->I, pr->Tt, (int64_t)c->Ip,
                       pr->syn_rot, pr->cell_of, pr->C, pr->Cp, idx, pr->Xb, pr->tb);
    return hipGetLastError();
}

hipError_t mlp_synth_counts(const emc_ctx *c, int n, float photons, uint64_t seed, hipStream_t st)
{
    const Proposal *pr = c->prop;
    if (c->opt.interp == EMC_INTERP_NEAREST)
        hipLaunchKernelGGL(k_synth_counts<EMC_INTERP_NEAREST>, dim3(n), dim3(256), 0, st, c->model, c->L, c->rotm,
                           pr->syn_rot, c->axyz, c->Pa, c->Pp, photons, seed, pr->syn_counts);
    else
        hipLaunchKernelGGL(k_synth_counts<EMC_INTERP_TRILINEAR>, dim3(n), dim3(256), 0, st, c->model, c->L, c->rotm,
                           pr->syn_rot, c->axyz, c->Pa, c->Pp, photons, seed, pr->syn_counts);
    return hipGetLastError();
}

}  // namespace emc
