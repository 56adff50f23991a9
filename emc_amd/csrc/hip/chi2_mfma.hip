// chi^2 likelihood kinds on the matrix cores (emc_cuda.cu:160-216, 242-276), with the optional
// per (rotation, pattern) intensity scaling (emc_cuda.cu:438-455, 516-531).
//
// The reference's D = sum_{valid p} a (W - K/phi)^2 (a = 1/(W + 0.02), 1 or 1/W^2 by kind) expands to
//   D = sum a W^2  -  2 ip sum (a W) K  +  ip^2 sum a K^2,      ip = 1/phi,
// two contractions over pixels of split-BF16 slice operands with the counts and the counts
// squared (both exact in BF16: K, and K^2 as hi + lo), plus a per-rotation constant.  The scaling's
// sums are two more contractions: sum K^2 over {W > 1e-10} (indicator operand, exact) and sum W K
// over the same pixels.  The FP32 products go through the large-tile kernel's store epilogue; one
// elementwise pass combines them into the logits (in double) and 1/phi.
#include "../emc_internal.h"

namespace emc {
namespace {

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// x -> NT BF16 terms whose sum carries ~8 NT significant bits of x.
template <int NT>
__device__ __forceinline__ void split(float x, bf16_t *t0, bf16_t *t1, bf16_t *t2, size_t o)
{
    const bf16_t h = f2bf(x);
    const float r1 = x - bf2f(h);
    const bf16_t m = f2bf(r1);
    t0[o] = h;
    t1[o] = m;
    if (NT == 3) t2[o] = f2bf(r1 - bf2f(m));
}

// One workgroup per chunk row r: the four operands from the FP32 slices (A0 = W, A1 = a; both 0 on
// invalid samples, masked pixels and padding rows) and Aw[r] = sum a W^2.
template <int NT>
__global__ __launch_bounds__(256) void k_chi2_operands(const float *__restrict__ A0, const float *__restrict__ A1,
                                                       int Pp, int64_t plane, bool scaling, bf16_t *__restrict__ Cx,
                                                       double *__restrict__ Aw)
{
    __shared__ double red[4];
    const int r = blockIdx.x;
    double acc = 0.0;
    bf16_t *t[4][3];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) t[p][q] = Cx + (int64_t)(p * NT + (q < NT ? q : 0)) * plane;
    for (int j = threadIdx.x; j < Pp; j += 256) {
        const size_t o = (size_t)r * Pp + j;
        const float W = A0[o], a = A1[o];
        acc += (double)a * W * W;
        split<NT>(a * W, t[0][0], t[0][1], t[0][2], o);
        split<NT>(a, t[1][0], t[1][1], t[1][2], o);
        if (scaling) {
            const bool in = W > 1.0e-10f;  // emc_cuda.cu:444 (a zero count adds nothing to either sum)
            t[2][0][o] = f2bf(in ? 1.0f : 0.0f);  // the indicator is exact in one term
            split<NT>(in ? W : 0.0f, t[3][0], t[3][1], t[3][2], o);
        }
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) Aw[r] = red[0] + red[1] + red[2] + red[3];
}

// R rows hold sum a W K on entry (rows [s0, s0 + cs), patterns [0, I)); G2 = sum a K^2,
// G3 = sum_{W>1e-10} K^2, G4 = sum_{W>1e-10} W K.  ip = G4 / G3 when both are > 0 (else 1).
__global__ void k_chi2_combine(float *__restrict__ R, const float *__restrict__ G2, const float *__restrict__ G3,
                               const float *__restrict__ G4, const double *__restrict__ Aw,
                               const float *__restrict__ bias, const float *__restrict__ rscale, int cs, int I,
                               int64_t ld, bool scaling, float *__restrict__ iphi)
{
    const int64_t n = (int64_t)cs * ld;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(t / ld), i = (int)(t % ld);
        if (i >= I) continue;
        double ip = 1.0;
        if (scaling) {
            const float a = G3[t], b = G4[t];
            const float f = (a > 0.0f && b > 0.0f) ? b / a : 1.0f;
            iphi[t] = f;
            ip = f;
        }
        const double D = Aw[r] - 2.0 * ip * (double)R[t] + ip * ip * (double)G2[t];
        R[t] = (float)((double)bias[r] - D * (double)rscale[r]);
    }
}

unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 16384); }

}  // namespace

hipError_t launch_logl_chi2_mfma(const emc_ctx *c, int s0, int cs, hipStream_t st)
{
    const int NT = c->la_terms;
    const int64_t plane = (int64_t)c->chunkp * c->Pp;
    const bool scaling = c->opt.scaling != 0;
    if (NT == 3)
        hipLaunchKernelGGL(k_chi2_operands<3>, dim3(c->chunkp), dim3(256), 0, st, c->A0, c->A1, c->Pp, plane, scaling,
                           c->Cx, c->Aw);
    else
        hipLaunchKernelGGL(k_chi2_operands<2>, dim3(c->chunkp), dim3(256), 0, st, c->A0, c->A1, c->Pp, plane, scaling,
                           c->Cx, c->Aw);
    hipError_t e = hipGetLastError();
    auto op = [&](int p, int q) { return c->Cx + (int64_t)(p * NT + q) * plane; };
    auto with_counts = [&](int p, float *C) {  // C = X_p . K^T (hi + lo counts above 256)
        hipError_t err = launch_gemm_store_mfma(c, NT, op(p, 0), op(p, 1), NT == 3 ? op(p, 2) : nullptr, c->K16, c->K8,
                                                C, false, st);
        if (err == hipSuccess && c->k_split)
            err = launch_gemm_store_mfma(c, NT, op(p, 0), op(p, 1), NT == 3 ? op(p, 2) : nullptr, c->K16lo, nullptr, C,
                                         true, st);
        return err;
    };
    // C = X_p . (K^2)^T with na terms of X_p; the low half of K^2 meets na - 1 of them (at least
    // one): the dropped product of the two lowest terms is below 2^-16 of the sum's terms.
    auto with_squares = [&](int p, int na, float *C) {
        hipError_t err = launch_gemm_store_mfma(c, na, op(p, 0), op(p, 1), na == 3 ? op(p, 2) : nullptr, c->K2h,
                                                nullptr, C, false, st);
        const int nl = na > 1 ? na - 1 : 1;
        if (err == hipSuccess && c->k2_split)
            err = launch_gemm_store_mfma(c, nl, op(p, 0), op(p, 1), nullptr, c->K2l, nullptr, C, true, st);
        return err;
    };
    float *R = c->R + (size_t)s0 * c->Ip;
    if (e == hipSuccess) e = with_counts(0, R);
    if (e == hipSuccess) e = with_squares(1, NT, c->G2);
    if (e == hipSuccess && scaling) e = with_squares(2, 1, c->G3);
    if (e == hipSuccess && scaling) e = with_counts(3, c->G4);
    if (e != hipSuccess) return e;
    const int64_t n = (int64_t)cs * c->Ip;
    hipLaunchKernelGGL(k_chi2_combine, dim3(grid_for(n)), dim3(256), 0, st, R, c->G2, c->G3, c->G4, c->Aw,
                       c->bias + s0, c->rscale + s0, cs, c->I, (int64_t)c->Ip, scaling,
                       scaling ? c->iphi + (size_t)s0 * c->Ip : nullptr);
    return hipGetLastError();
}

}  // namespace emc
