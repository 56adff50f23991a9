// Expand: quaternion slicing of the 3D model (emc_cuda.cu:102-157) and the per-rotation
// quantities the likelihood needs.
//
// Layout: active (mask != 0) pixels are compacted to j in [0, Pa) with coordinates axyz[3][Pp];
// slices are written row-per-rotation [chunkp][Pp] so that both contractions read them with
// the reduction index contiguous.  One workgroup owns one rotation row: the rotation matrix is
// wave-uniform, coordinate reads are coalesced, and model reads are gathers that hit L2
// (the 64^3 model is 1 MiB).
#include <algorithm>

#include "../emc_internal.h"
#include "geom.h"

namespace emc {
namespace {

// rotations per workgroup in the statistics pass: 8 keeps 4 waves per SIMD (16 needed 210 VGPRs,
// 2 waves per SIMD: 0.77 instead of 0.61 ms at 64^3 / 25,680 rotations)
constexpr int kStatsG = 8;

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ void k_rotmat(const float *__restrict__ q, const float *__restrict__ w, int S, int Sp,
                         float *__restrict__ rotm, float *__restrict__ logw)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= Sp) return;
    float m[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    float lw = 0.0f;
    if (s < S) {
        rotmat(q[4 * s], q[4 * s + 1], q[4 * s + 2], q[4 * s + 3], m);
        lw = logf(w[s]);
    }
    for (int k = 0; k < 12; ++k) rotm[12 * (size_t)s + k] = m[k];
    logw[s] = lw;
}

template <typename T>
__global__ void k_fill(T *__restrict__ p, T v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Per-rotation slice sum over V_s = {active j : W > 0}, |V_s|, and per pixel the workgroup's sum of
// log W and number of rotations with a valid sample (colp[block][j] = (sum, count); summed over
// the blocks in a fixed order by k_colsum_*, so the column sums are deterministic).
// kStatsG rotations per workgroup; each thread walks pixels j = tid, tid+256, ... and keeps the
// kStatsG partial sums in registers.  With lw, the pass also keeps log W ([Sp][Pp], -inf where
// W <= 0) for the operand split, which then needs no second gather from the model.
// modelq[v] = (model[v], model[v + 1], model[v + L], model[v + L + 1]), 0 past an edge (the trilinear
// statistics pass reads the 8 corners of a sample as two 16-byte loads).
__global__ void k_quad_model(const float *__restrict__ model, int L, float4 *__restrict__ modelq)
{
    const int64_t n = (int64_t)L * L * L;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const bool xe = (int)(v % L) + 1 < L, ye = (int)((v / L) % L) + 1 < L;
        modelq[v] = make_float4(model[v], xe ? model[v + 1] : 0.0f, ye ? model[v + L] : 0.0f,
                                xe && ye ? model[v + L + 1] : 0.0f);
    }
}

template <int INTERP>
__global__ __launch_bounds__(256) void k_slice_stats(const float *__restrict__ model,
                                                     const float4 *__restrict__ modelq, int L,
                                                     const float *__restrict__ rotm, int S,
                                                     const float *__restrict__ axyz, int Pa, int Pp,
                                                     double *__restrict__ sumW, int32_t *__restrict__ ns,
                                                     float2 *__restrict__ colp, float *__restrict__ lw)
{
    __shared__ float rm[kStatsG][9];
    __shared__ double red_w[kStatsG][256];
    __shared__ int red_n[kStatsG][256];
    static_assert(kStatsG * 32 == 256, "32 lanes per rotation in the final sums");
    const int s0 = blockIdx.x * kStatsG;
    const int tid = threadIdx.x;
    if (tid < kStatsG * 9) {
        const int g = tid / 9, k = tid % 9;
        rm[g][k] = (s0 + g < S) ? rotm[12 * (size_t)(s0 + g) + k] : 0.0f;
    }
    __syncthreads();
    const int ng = min(kStatsG, S - s0);
    float sw[kStatsG];
    int cnt[kStatsG];
#pragma unroll
    for (int g = 0; g < kStatsG; ++g) { sw[g] = 0.0f; cnt[g] = 0; }
    for (int j = tid; j < Pa; j += 256) {
        const float X = axyz[j], Y = axyz[Pp + j], Z = axyz[2 * Pp + j];
        float col = 0.0f;
        int colv = 0;
#pragma unroll
        for (int g = 0; g < kStatsG; ++g) {
            if (g < ng) {
                const float W = INTERP == EMC_INTERP_TRILINEAR ? sample_quad(modelq, L, rm[g], X, Y, Z)
                                                               : sample<INTERP>(model, L, rm[g], X, Y, Z);
                float l = -INFINITY;
                if (W > 0.0f) {
                    l = logf(W);
                    sw[g] += W;
                    cnt[g] += 1;
                    col += l;
                    colv += 1;
                }
                if (lw) lw[(size_t)(s0 + g) * Pp + j] = l;
            }
        }
        if (colp) colp[(size_t)blockIdx.x * Pp + j] = make_float2(col, (float)colv);
    }
    // per-rotation sums over the workgroup's 256 threads through LDS, in a fixed order: 32 lanes per
    // rotation add 8 values each, then five shuffle steps
    const int lane = tid & 63;
#pragma unroll
    for (int g = 0; g < kStatsG; ++g) {
        red_w[g][tid] = (double)sw[g];
        red_n[g][tid] = cnt[g];
    }
    __syncthreads();
    {
        const int g = tid >> 5, part = lane & 31;
        double a = 0.0;
        int b = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            a += red_w[g][part + 32 * q];
            b += red_n[g][part + 32 * q];
        }
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) {
            a += __shfl_xor(a, o, 64);
            b += __shfl_xor(b, o, 64);
        }
        if (part == 0 && g < ng) {
            sumW[s0 + g] = a;
            ns[s0 + g] = b;
        }
    }
}

// Column sums of the statistics pass, in two fixed-order steps: part r of the blocks
// [r * per, (r + 1) * per) per pixel, then the kColParts parts in order.
constexpr int kColParts = 32;

__global__ __launch_bounds__(256) void k_colsum_part(const float2 *__restrict__ colp, int nblk, int Pa, int Pp,
                                                     double2 *__restrict__ cpart)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= Pa) return;
    const int r = blockIdx.y, per = (nblk + kColParts - 1) / kColParts;
    const int b0 = r * per, b1 = min(nblk, b0 + per);
    double a = 0.0, n = 0.0;
    int b = b0;
    for (; b + 4 <= b1; b += 4) {  // four independent loads, added in order
        const float2 v0 = colp[(size_t)b * Pp + j], v1 = colp[(size_t)(b + 1) * Pp + j];
        const float2 v2 = colp[(size_t)(b + 2) * Pp + j], v3 = colp[(size_t)(b + 3) * Pp + j];
        a += v0.x; n += v0.y;
        a += v1.x; n += v1.y;
        a += v2.x; n += v2.y;
        a += v3.x; n += v3.y;
    }
    for (; b < b1; ++b) {
        const float2 v = colp[(size_t)b * Pp + j];
        a += v.x;
        n += v.y;
    }
    cpart[(size_t)r * Pp + j] = make_double2(a, n);
}

__global__ __launch_bounds__(256) void k_colsum_final(const double2 *__restrict__ cpart, int Pa, int Pp,
                                                      double *__restrict__ colsum)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= Pp) return;
    double a = 0.0, n = 0.0;
    if (j < Pa)
        for (int r = 0; r < kColParts; ++r) {
            const double2 v = cpart[(size_t)r * Pp + j];
            a += v.x;
            n += v.y;
        }
    colsum[2 * j] = a;
    colsum[2 * j + 1] = n;
}

// FP32 slice rows for the plain contraction.  KIND: TRUE_POISSON -> A0 = log(phi0 W) (0 if
// invalid; log phi0 is read from the device scalars written by k_bias);
// chi^2 kinds -> A0 = W, A1 = pixel weight a (0 if invalid):
//   poisson (reference chi^2) a = 1/(W+0.02), absolute a = 1, relative a = 1/W^2.
template <int INTERP, int KIND>
__global__ __launch_bounds__(256) void k_slice_f32(const float *__restrict__ model, int L,
                                                   const float *__restrict__ rotm, int s0, int cs,
                                                   const float *__restrict__ axyz, int Pa, int Pp,
                                                   const double *__restrict__ scal,
                                                   float *__restrict__ A0, float *__restrict__ A1)
{
    const int r = blockIdx.x;
    const bool row_ok = r < cs;
    const float logphi0 = KIND == EMC_DIFF_TRUE_POISSON ? (float)scal[SC_LOGPHI0] : 0.0f;
    float m[9];
    if (row_ok) load_rot(rotm, s0 + r, m);
    for (int j = threadIdx.x; j < Pp; j += 256) {
        float a0 = 0.0f, a1 = 0.0f;
        if (row_ok && j < Pa) {
            const float W = sample<INTERP>(model, L, m, axyz[j], axyz[Pp + j], axyz[2 * Pp + j]);
            if (W > 0.0f) {
                if (KIND == EMC_DIFF_TRUE_POISSON) {
                    a0 = logf(W) + logphi0;
                } else {
                    a0 = W;
                    a1 = KIND == EMC_DIFF_POISSON ? 1.0f / (W + 0.02f)
                                                  : (KIND == EMC_DIFF_ABSOLUTE ? 1.0f : 1.0f / (W * W));
                }
            }
        }
        A0[(size_t)r * Pp + j] = a0;
        if (KIND != EMC_DIFF_TRUE_POISSON) A1[(size_t)r * Pp + j] = a1;
    }
}

// Centred log-slice x = (W > 0 ? log(phi0 W) : 0) - colmean[j] split into NT BF16 terms
// (hi, [mid,] lo) whose sum carries ~8*NT significant bits of x.
template <int INTERP, int NT>
__global__ __launch_bounds__(256) void k_slice_bf16(const float *__restrict__ model, int L,
                                                    const float *__restrict__ rotm, int s0, int cs,
                                                    const float *__restrict__ axyz, int Pa, int Pp,
                                                    const float *__restrict__ colmean,
                                                    const double *__restrict__ scal,
                                                    bf16_t *__restrict__ Lh, bf16_t *__restrict__ Lm,
                                                    bf16_t *__restrict__ Ll)
{
    const int r = blockIdx.x;
    const bool row_ok = r < cs;
    const float logphi0 = (float)scal[SC_LOGPHI0];
    float m[9];
    if (row_ok) load_rot(rotm, s0 + r, m);
    for (int j = threadIdx.x; j < Pp; j += 256) {
        float x = 0.0f;
        if (row_ok && j < Pa) {
            const float W = sample<INTERP>(model, L, m, axyz[j], axyz[Pp + j], axyz[2 * Pp + j]);
            x = (W > 0.0f ? logf(W) + logphi0 : 0.0f) - colmean[j];
        }
        const bf16_t h = f2bf(x);
        const float r1 = x - bf2f(h);
        const bf16_t mi = f2bf(r1);
        const size_t o = (size_t)r * Pp + j;
        Lh[o] = h;
        if (NT == 2) {
            Ll[o] = mi;
        } else {
            Lm[o] = mi;
            Ll[o] = f2bf(r1 - bf2f(mi));
        }
    }
}

// The same operand split from the log W kept by the statistics pass: x = (log W + log phi0, or 0
// where W <= 0) - colmean[j], the same float operations as k_slice_bf16, so the terms are
// bit-identical.  Four pixels per thread and kSplitRows rows per workgroup, whose loads are issued
// together (the column means are read once for all of them).  Pp is a multiple of 128.
constexpr int kSplitRows = 8;

template <int NT>
__global__ __launch_bounds__(256) void k_split_bf16(const float *__restrict__ lw, int s0, int cs, int Pa, int Pp,
                                                    const float *__restrict__ colmean,
                                                    const double *__restrict__ scal, bf16_t *__restrict__ Lh,
                                                    bf16_t *__restrict__ Lm, bf16_t *__restrict__ Ll)
{
    const int r0 = blockIdx.y * kSplitRows;
    const int j0 = (blockIdx.x * 256 + threadIdx.x) * 4;
    if (j0 >= Pp) return;
    const float logphi0 = (float)scal[SC_LOGPHI0];
    float cm[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) cm[q] = j0 + q < Pa ? colmean[j0 + q] : 0.0f;
    float4 v[kSplitRows];
#pragma unroll
    for (int rr = 0; rr < kSplitRows; ++rr) {
        v[rr] = make_float4(-INFINITY, -INFINITY, -INFINITY, -INFINITY);
        if (r0 + rr < cs) v[rr] = *reinterpret_cast<const float4 *>(lw + (size_t)(s0 + r0 + rr) * Pp + j0);
    }
    auto pack = [](const uint16_t *w) {
        return make_uint2((uint32_t)w[0] | ((uint32_t)w[1] << 16), (uint32_t)w[2] | ((uint32_t)w[3] << 16));
    };
#pragma unroll
    for (int rr = 0; rr < kSplitRows; ++rr) {
        const int r = r0 + rr;  // rows of the padded chunk past cs are written as zeros
        const float l[4] = {v[rr].x, v[rr].y, v[rr].z, v[rr].w};
        uint16_t h[4], mi[4], lo[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = j0 + q;
            float x = 0.0f;
            if (r < cs && j < Pa) x = (l[q] > -INFINITY ? l[q] + logphi0 : 0.0f) - cm[q];
            h[q] = f2bf(x);
            const float r1 = x - bf2f(h[q]);
            mi[q] = f2bf(r1);
            lo[q] = f2bf(r1 - bf2f(mi[q]));
        }
        const size_t o = (size_t)r * Pp + j0;
        *reinterpret_cast<uint2 *>(Lh + o) = pack(h);
        if (NT == 2) {
            *reinterpret_cast<uint2 *>(Ll + o) = pack(mi);
        } else {
            *reinterpret_cast<uint2 *>(Lm + o) = pack(mi);
            *reinterpret_cast<uint2 *>(Ll + o) = pack(lo);
        }
    }
}

template <int INTERP>
__global__ __launch_bounds__(256) void k_expand(const float *__restrict__ model, int L,
                                                const float *__restrict__ rotm, int s0,
                                                const float *__restrict__ coords, int P,
                                                float *__restrict__ out)
{
    const int r = blockIdx.x;
    float m[9];
    load_rot(rotm, s0 + r, m);
    for (int p = threadIdx.x; p < P; p += 256)
        out[(size_t)r * P + p] = sample<INTERP>(model, L, m, coords[p], coords[P + p], coords[2 * P + p]);
}

// Fluence, per-rotation logit terms and pixel column means.  Every workgroup forms the total
// sum_s sumW_s itself (the same fixed order in each, so the same value) and writes its share of the
// rotations and pixels; workgroup 0 writes the scalars.
//   TRUE_POISSON: phi0 = fluence or (sumK/I)/(sum_s sumW_s / S);
//                 bias_s = log w_s - phi0 sumW_s + B0, B0 = phi0 * mean_s sumW_s (a global
//                 shift that keeps the stored logits small; it is returned through cst).
//   chi^2 kinds:  bias_s = log w_s, rscale_s = 1/(2 n_s sigma^2) (0 when n_s = 0).
__global__ __launch_bounds__(1024) void k_bias(const double *__restrict__ sumW, const int32_t *__restrict__ ns,
                                               const float *__restrict__ logw, int S, int Sp, int I,
                                               double sumK_total, double fluence, int diff, float sigma,
                                               const double *__restrict__ colsum, int Pa, int Pp,
                                               float *__restrict__ bias, float *__restrict__ rscale,
                                               float *__restrict__ colmean, double *__restrict__ scal)
{
    __shared__ double red[16];
    const int tid = threadIdx.x;
    // eight partial sums per thread (eight loads in flight), added in a fixed order
    double p[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int s8 = tid;
    for (; s8 + 7 * 1024 < S; s8 += 8 * 1024)
#pragma unroll
        for (int q = 0; q < 8; ++q) p[q] += sumW[s8 + q * 1024];
#pragma unroll
    for (int q = 0; q < 8; ++q)
        if (s8 + q * 1024 < S) p[q] += sumW[s8 + q * 1024];
    double acc = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
    acc = wave_sum(acc);
    if ((tid & 63) == 0) red[tid >> 6] = acc;
    __syncthreads();
    double SW = 0.0;
    for (int k = 0; k < 16; ++k) SW += red[k];
    double phi0 = 1.0;
    if (diff == EMC_DIFF_TRUE_POISSON) {
        if (fluence > 0.0) phi0 = fluence;
        else if (SW > 0.0 && I > 0 && sumK_total > 0.0) phi0 = (sumK_total / I) / (SW / S);  // else undefined: 1
    }
    const double B0 = (diff == EMC_DIFF_TRUE_POISSON && S > 0) ? phi0 * SW / S : 0.0;
    if (blockIdx.x == 0 && tid == 0) {
        scal[SC_PHI0] = phi0;
        scal[SC_LOGPHI0] = log(phi0);
        scal[SC_SUMW] = SW;
        scal[SC_B0] = B0;
    }
    const double inv2s2 = 1.0 / (2.0 * (double)sigma * (double)sigma);
    const int t0 = blockIdx.x * 1024 + tid, stride = gridDim.x * 1024;
#pragma unroll 4
    for (int s = t0; s < Sp; s += stride) {
        float b = 0.0f, rs = 0.0f;
        if (s < S) {
            if (diff == EMC_DIFF_TRUE_POISSON) b = (float)((double)logw[s] - phi0 * sumW[s] + B0);
            else {
                b = logw[s];
                rs = ns[s] > 0 ? (float)(inv2s2 / (double)ns[s]) : 0.0f;
            }
        }
        bias[s] = b;
        rscale[s] = rs;
    }
    // column mean of log(phi0 W) over rotations (invalid samples count as 0)
    const double lp = log(phi0);
    for (int j = t0; j < Pp; j += stride)
        colmean[j] = (colsum && j < Pa && S > 0) ? (float)((colsum[2 * j] + colsum[2 * j + 1] * lp) / S) : 0.0f;
}

// cst_i = sum_j K_ij colmean_j - B0   (FP64; one wavefront per pattern, four per workgroup)
__global__ __launch_bounds__(256) void k_pattern_const(const float *__restrict__ K32,
                                                       const bf16_t *__restrict__ K16,
                                                       const bf16_t *__restrict__ K16lo,
                                                       const uint8_t *__restrict__ K8, int I, int Pp,
                                                       const float *__restrict__ colmean,
                                                       const double *__restrict__ sumK,
                                                       const double *__restrict__ scal, bool centred,
                                                       double *__restrict__ cst)
{
    const int lane = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= I) return;  // a whole wavefront
    double acc = 0.0;
    if (centred && K8) {  // byte counts (exact): 16 per lane and 16-byte load (Pp % 16 == 0)
        for (int j = lane * 16; j < Pp; j += 64 * 16) {
            const uint4 v = *reinterpret_cast<const uint4 *>(K8 + (size_t)i * Pp + j);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 m = *reinterpret_cast<const float4 *>(colmean + j + 4 * q);
                acc = fma((double)(w[q] & 0xffu), (double)m.x, acc);
                acc = fma((double)((w[q] >> 8) & 0xffu), (double)m.y, acc);
                acc = fma((double)((w[q] >> 16) & 0xffu), (double)m.z, acc);
                acc = fma((double)(w[q] >> 24), (double)m.w, acc);
            }
        }
    } else if (centred) {
        for (int j = lane; j < Pp; j += 64) {
            float k;
            if (K32) k = K32[(size_t)i * Pp + j];
            else {
                k = bf2f(K16[(size_t)i * Pp + j]);
                if (K16lo) k += bf2f(K16lo[(size_t)i * Pp + j]);
            }
            acc += (double)k * (double)colmean[j];
        }
    }
    acc = wave_sum(acc);
    if (lane == 0) cst[i] = acc - scal[SC_B0];
}

}  // namespace

hipError_t launch_rotmat(const float *q, const float *w, int S, int Sp, float *rotm, float *logw,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_rotmat, dim3((Sp + 255) / 256), dim3(256), 0, st, q, w, S, Sp, rotm, logw);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float *p, float v, int64_t n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill<float>, dim3((unsigned)blocks), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_f64(double *p, double v, int64_t n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)blocks), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill<int32_t>, dim3((unsigned)blocks), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

int64_t colp_len(int S, int Pp) { return (int64_t)((S + kStatsG - 1) / kStatsG) * Pp; }
int64_t colpart_len(int Pp) { return (int64_t)kColParts * Pp; }

hipError_t launch_slice_stats(const emc_ctx *c, bool want_colsum, hipStream_t st)
{
    const int nblk = (c->S + kStatsG - 1) / kStatsG;
    float2 *cp = want_colsum ? c->colp : nullptr;
    if (c->opt.interp == EMC_INTERP_NEAREST) {
        hipLaunchKernelGGL(k_slice_stats<EMC_INTERP_NEAREST>, dim3(nblk), dim3(256), 0, st, c->model, nullptr, c->L,
                           c->rotm, c->S, c->axyz, c->Pa, c->Pp, c->sumW, c->ns, cp, c->lw32);
    } else {
        const int64_t n3 = (int64_t)c->L * c->L * c->L;
        hipLaunchKernelGGL(k_quad_model, dim3((unsigned)std::min<int64_t>((n3 + 255) / 256, 4096)), dim3(256), 0, st,
                           c->model, c->L, c->modelq);
        hipLaunchKernelGGL(k_slice_stats<EMC_INTERP_TRILINEAR>, dim3(nblk), dim3(256), 0, st, c->model, c->modelq,
                           c->L, c->rotm, c->S, c->axyz, c->Pa, c->Pp, c->sumW, c->ns, cp, c->lw32);
    }
    if (want_colsum) {
        hipLaunchKernelGGL(k_colsum_part, dim3((c->Pa + 255) / 256, kColParts), dim3(256), 0, st, c->colp, nblk, c->Pa,
                           c->Pp, c->colpart);
        hipLaunchKernelGGL(k_colsum_final, dim3((c->Pp + 255) / 256), dim3(256), 0, st, c->colpart, c->Pa, c->Pp,
                           c->colsum);
    }
    return hipGetLastError();
}

hipError_t launch_bias(const emc_ctx *c, float sigma, hipStream_t st)
{
    const int nb = std::max(1, (std::max(c->Sp, c->Pp) + 1023) / 1024);
    hipLaunchKernelGGL(k_bias, dim3(nb), dim3(1024), 0, st, c->sumW, c->ns, c->logw, c->S, c->Sp, c->I,
                       c->sumK_total, c->opt.fluence, c->opt.diff, sigma, c->mfma ? c->colsum : nullptr,
                       c->Pa, c->Pp, c->bias, c->rscale, c->colmean, c->scal);
    return hipGetLastError();
}

hipError_t launch_pattern_const(const emc_ctx *c, bool centred, hipStream_t st)
{
    hipLaunchKernelGGL(k_pattern_const, dim3((c->I + 3) / 4), dim3(256), 0, st, c->K32, c->K16, c->k_split ? c->K16lo : nullptr, c->K8, c->I,
                       c->Pp, c->colmean, c->sumK, c->scal, centred, c->cst);
    return hipGetLastError();
}

template <int INTERP>
static hipError_t slice_f32_interp(const emc_ctx *c, int s0, int cs, hipStream_t st)
{
    const dim3 grid(c->chunkp), blk(256);
    switch (c->opt.diff) {
    case EMC_DIFF_TRUE_POISSON:
        hipLaunchKernelGGL((k_slice_f32<INTERP, EMC_DIFF_TRUE_POISSON>), grid, blk, 0, st, c->model, c->L,
                           c->rotm, s0, cs, c->axyz, c->Pa, c->Pp, c->scal, c->A0, c->A1);
        break;
    case EMC_DIFF_POISSON:
        hipLaunchKernelGGL((k_slice_f32<INTERP, EMC_DIFF_POISSON>), grid, blk, 0, st, c->model, c->L,
                           c->rotm, s0, cs, c->axyz, c->Pa, c->Pp, c->scal, c->A0, c->A1);
        break;
    case EMC_DIFF_ABSOLUTE:
        hipLaunchKernelGGL((k_slice_f32<INTERP, EMC_DIFF_ABSOLUTE>), grid, blk, 0, st, c->model, c->L,
                           c->rotm, s0, cs, c->axyz, c->Pa, c->Pp, c->scal, c->A0, c->A1);
        break;
    default:
        hipLaunchKernelGGL((k_slice_f32<INTERP, EMC_DIFF_RELATIVE>), grid, blk, 0, st, c->model, c->L,
                           c->rotm, s0, cs, c->axyz, c->Pa, c->Pp, c->scal, c->A0, c->A1);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_slice_f32(const emc_ctx *c, int s0, int cs, hipStream_t st)
{
    return c->opt.interp == EMC_INTERP_NEAREST ? slice_f32_interp<EMC_INTERP_NEAREST>(c, s0, cs, st)
                                               : slice_f32_interp<EMC_INTERP_TRILINEAR>(c, s0, cs, st);
}

hipError_t launch_slice_bf16(const emc_ctx *c, int s0, int cs, hipStream_t st)
{
    if (c->lw32) {
        const dim3 g((c->Pp / 4 + 255) / 256, (c->chunkp + kSplitRows - 1) / kSplitRows);
        if (c->la_terms == 3)
            hipLaunchKernelGGL(k_split_bf16<3>, g, dim3(256), 0, st, c->lw32, s0, cs, c->Pa, c->Pp, c->colmean, c->scal,
                               c->Lh, c->Lm, c->Ll);
        else
            hipLaunchKernelGGL(k_split_bf16<2>, g, dim3(256), 0, st, c->lw32, s0, cs, c->Pa, c->Pp, c->colmean, c->scal,
                               c->Lh, c->Lm, c->Ll);
        return hipGetLastError();
    }
    const dim3 grid(c->chunkp), blk(256);
    const bool nn = c->opt.interp == EMC_INTERP_NEAREST;
    if (c->la_terms == 3) {
        if (nn)
            hipLaunchKernelGGL((k_slice_bf16<EMC_INTERP_NEAREST, 3>), grid, blk, 0, st, c->model, c->L, c->rotm,
                               s0, cs, c->axyz, c->Pa, c->Pp, c->colmean, c->scal, c->Lh, c->Lm, c->Ll);
        else
            hipLaunchKernelGGL((k_slice_bf16<EMC_INTERP_TRILINEAR, 3>), grid, blk, 0, st, c->model, c->L, c->rotm,
                               s0, cs, c->axyz, c->Pa, c->Pp, c->colmean, c->scal, c->Lh, c->Lm, c->Ll);
    } else {
        if (nn)
            hipLaunchKernelGGL((k_slice_bf16<EMC_INTERP_NEAREST, 2>), grid, blk, 0, st, c->model, c->L, c->rotm,
                               s0, cs, c->axyz, c->Pa, c->Pp, c->colmean, c->scal, c->Lh, c->Lm, c->Ll);
        else
            hipLaunchKernelGGL((k_slice_bf16<EMC_INTERP_TRILINEAR, 2>), grid, blk, 0, st, c->model, c->L, c->rotm,
                               s0, cs, c->axyz, c->Pa, c->Pp, c->colmean, c->scal, c->Lh, c->Lm, c->Ll);
    }
    return hipGetLastError();
}

hipError_t launch_expand(const emc_ctx *c, int s0, int cs, float *out, hipStream_t st)
{
    if (cs <= 0) return hipSuccess;
    if (c->opt.interp == EMC_INTERP_NEAREST)
        hipLaunchKernelGGL(k_expand<EMC_INTERP_NEAREST>, dim3(cs), dim3(256), 0, st, c->model, c->L, c->rotm, s0,
                           c->coords, c->P, out);
    else
        hipLaunchKernelGGL(k_expand<EMC_INTERP_TRILINEAR>, dim3(cs), dim3(256), 0, st, c->model, c->L, c->rotm,
                           s0, c->coords, c->P, out);
    return hipGetLastError();
}

}  // namespace emc
