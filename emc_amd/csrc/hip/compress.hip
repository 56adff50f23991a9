// Compress: divide by weight (emc_cuda.cu:758-786), normalisation to mean 1 (emc_cuda.cu:639-661,
// 788-798; FP64 grid-wide sums instead of the reference's single-block FP32 reduction) and the
// masked Gaussian blur (the intended math of emc_cuda.cu:1404-1492), in that order.
//
// The reference averages the model over its positive voxels.  Which voxels are positive after the
// divide depends on where the responsibilities underflow (a voxel whose every contribution is a
// P below the float range is 0 in FP32 and positive in FP64: 6 % of the voxels at 32^3), and after
// the blur on the sign of rounding-level ringing.  Here the mean is taken over the voxels that
// received weight (Wt > 0, zeros included) of the divided model: the same set in every precision
// and at every summation order, and the same mean wherever the support holds no zero.  The
// initial model (no weights yet) is averaged over its positive voxels as in the reference.
//
// Blur: the Fourier multiplier exp(-2 pi^2 sigma^2 |d|^2 / L^2), d = min(k, L-k) per axis, is
// separable and real-symmetric, so it equals three periodic 1-D convolutions with taps
// g[d] = (1/L) sum_k h(k) cos(2 pi k d / L) (precomputed on the host in FP64).  At L = 64 this is
// 3 * L^4 = 5e7 FMAs on an L2-resident volume, cheaper than a 3-D FFT round trip.
#include "../emc_internal.h"

namespace emc {
namespace {

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ void k_divide(const float *__restrict__ Mnum, const float *__restrict__ Wt, float *__restrict__ M,
                         int64_t n)
{
    const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= n) return;
    const float w = Wt[v];
    M[v] = w > 0.0f ? Mnum[v] / w : -1.0f;
}

// One axis of the periodic convolution.  PHASE 0 reads the model with negatives as 0; PHASE 2
// writes the result back into the model, where a negative (missing) voxel keeps its value.
template <int PHASE>
__global__ __launch_bounds__(256) void k_blur_axis(const float *__restrict__ in, float *__restrict__ out,
                                                   float *__restrict__ model, const float *__restrict__ taps, int L,
                                                   int axis)
{
    extern __shared__ float g[];
    for (int k = threadIdx.x; k < L; k += blockDim.x) g[k] = taps[k];
    __syncthreads();
    const int64_t N3 = (int64_t)L * L * L;
    const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= N3) return;
    const int x = (int)(v % L), y = (int)((v / L) % L), z = (int)(v / ((int64_t)L * L));
    const int coord = axis == 0 ? x : (axis == 1 ? y : z);
    const int64_t stride = axis == 0 ? 1 : (axis == 1 ? L : (int64_t)L * L);
    const int64_t base = v - (int64_t)coord * stride;
    float acc = 0.0f;
    for (int t = 0; t < L; ++t) {
        int d = coord - t;
        d += d < 0 ? L : 0;
        float a = in[base + (int64_t)t * stride];
        if (PHASE == 0) a = a < 0.0f ? 0.0f : a;
        acc = __builtin_fmaf(g[d], a, acc);
    }
    if (PHASE == 2) {
        const float m = model[v];
        model[v] = m < 0.0f ? m : acc;
    }
    else out[v] = acc;
}

// Per-block FP64 sum and count over the averaged voxels (those with Wt > 0 when Wt is given, else
// the positive ones, emc_cuda.cu:639-661), the max (emc_cuda.cu:629-637) and the sum of Wt.
__global__ __launch_bounds__(256) void k_model_partial(const float *__restrict__ M, const float *__restrict__ Wt,
                                                       int64_t n, double *__restrict__ part)
{
    __shared__ double rs[4], rc[4], rm[4], rw[4];
    double s = 0.0, cnt = 0.0, mx = -1e300, ws = 0.0;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const float x = M[v];
        const float w = Wt ? Wt[v] : 0.0f;
        if (Wt ? w > 0.0f : x > 0.0f) { s += x; cnt += 1.0; }
        mx = fmax(mx, (double)x);
        ws += w;
    }
    s = wave_sum(s);
    cnt = wave_sum(cnt);
    ws = wave_sum(ws);
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) {
        rs[threadIdx.x >> 6] = s; rc[threadIdx.x >> 6] = cnt; rm[threadIdx.x >> 6] = mx; rw[threadIdx.x >> 6] = ws;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x + 0] = rs[0] + rs[1] + rs[2] + rs[3];
        part[4 * blockIdx.x + 1] = rc[0] + rc[1] + rc[2] + rc[3];
        part[4 * blockIdx.x + 2] = fmax(fmax(rm[0], rm[1]), fmax(rm[2], rm[3]));
        part[4 * blockIdx.x + 3] = rw[0] + rw[1] + rw[2] + rw[3];
    }
}

__global__ __launch_bounds__(256) void k_model_final(const double *__restrict__ part, int nb, double *__restrict__ scal,
                                                     bool with_wt)
{
    __shared__ double rs[4], rc[4], rm[4], rw[4];
    double s = 0.0, cnt = 0.0, mx = -1e300, ws = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) {
        s += part[4 * b];
        cnt += part[4 * b + 1];
        mx = fmax(mx, part[4 * b + 2]);
        ws += part[4 * b + 3];
    }
    s = wave_sum(s);
    cnt = wave_sum(cnt);
    ws = wave_sum(ws);
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) {
        rs[threadIdx.x >> 6] = s; rc[threadIdx.x >> 6] = cnt; rm[threadIdx.x >> 6] = mx; rw[threadIdx.x >> 6] = ws;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double S = rs[0] + rs[1] + rs[2] + rs[3], C = rc[0] + rc[1] + rc[2] + rc[3];
        const double avg = C > 0.0 ? S / C : 0.0;
        scal[SC_AVG] = avg;
        scal[SC_COUNT] = C;
        scal[SC_INV] = avg > 0.0 ? 1.0 / avg : 1.0;
        scal[SC_MAX] = fmax(fmax(rm[0], rm[1]), fmax(rm[2], rm[3]));
        if (with_wt) scal[SC_WSUM] = rw[0] + rw[1] + rw[2] + rw[3];
        if (!(avg > 0.0) || !isfinite(avg)) scal[SC_NONFINITE] = 2.0;
    }
}

// M *= 1/avg on every voxel, the negative (missing) ones included (emc_cuda.cu:795).
__global__ void k_scale(float *__restrict__ M, int64_t n, const double *__restrict__ scal)
{
    const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= n) return;
    M[v] = (float)((double)M[v] * scal[SC_INV]);
}

constexpr int kPartBlocks = 256;

// Blur taps g[d] = (1/L) sum_k exp(-2 pi^2 sigma^2 min(k, L-k)^2 / L^2) cos(2 pi k d / L), FP64.
__global__ void k_blur_taps(int L, double sigma, float *__restrict__ taps)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= L) return;
    double acc = 0.0;
    for (int k = 0; k < L; ++k) {
        const int kk = min(k, L - k);
        acc += exp(-2.0 * M_PI * M_PI * sigma * sigma * (double)kk * kk / ((double)L * L)) *
               cos(2.0 * M_PI * (double)k * d / L);
    }
    taps[d] = (float)(acc / L);
}

}  // namespace

hipError_t launch_blur_taps(int L, double sigma, float *taps, hipStream_t st)
{
    hipLaunchKernelGGL(k_blur_taps, dim3((L + 63) / 64), dim3(64), 0, st, L, sigma, taps);
    return hipGetLastError();
}

// The steps of one compress on explicit buffers: divide, normalise over the voxels with weight,
// blur (t1 and t2: two L^3 volumes of scratch).
hipError_t launch_compress_buffers(int L, bool blur, const float *Mnum, const float *Wt, float *model,
                                   const float *taps, float *t1, float *t2, double *red, double *scal, hipStream_t st)
{
    const int64_t N3 = (int64_t)L * L * L;
    const dim3 grid((unsigned)((N3 + 255) / 256)), blk(256);
    hipLaunchKernelGGL(k_divide, grid, blk, 0, st, Mnum, Wt, model, N3);
    hipLaunchKernelGGL(k_model_partial, dim3(kPartBlocks), blk, 0, st, model, Wt, N3, red);
    hipLaunchKernelGGL(k_model_final, dim3(1), blk, 0, st, red, kPartBlocks, scal, true);
    hipLaunchKernelGGL(k_scale, grid, blk, 0, st, model, N3, scal);
    if (blur) {
        const size_t sh = sizeof(float) * L;
        hipLaunchKernelGGL(k_blur_axis<0>, grid, blk, sh, st, model, t1, nullptr, taps, L, 0);
        hipLaunchKernelGGL(k_blur_axis<1>, grid, blk, sh, st, t1, t2, nullptr, taps, L, 1);
        hipLaunchKernelGGL(k_blur_axis<2>, grid, blk, sh, st, t2, nullptr, model, taps, L, 2);
    }
    return hipGetLastError();
}

hipError_t launch_normalize(const emc_ctx *c, bool apply, bool support, hipStream_t st)
{
    static_assert(4 * kPartBlocks <= kRedLen, "model statistics partials");
    hipLaunchKernelGGL(k_model_partial, dim3(kPartBlocks), dim3(256), 0, st, c->model, support ? c->Wt : nullptr,
                       c->N3, c->red);
    hipLaunchKernelGGL(k_model_final, dim3(1), dim3(256), 0, st, c->red, kPartBlocks, c->scal, support);
    if (apply)
        hipLaunchKernelGGL(k_scale, dim3((unsigned)((c->N3 + 255) / 256)), dim3(256), 0, st, c->model, c->N3,
                           c->scal);
    return hipGetLastError();
}

}  // namespace emc
