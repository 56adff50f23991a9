// Plain FP32 contractions (the B-FP32 path): direct FP32 on the vector ALU.
//
//   logL   : R[s,i] = sum_j f(A0[s,j], A1[s,j], K[i,j]) then the epilogue
//              TRUE_POISSON  f = A0*K              R = acc + bias_s            (emc_cuda.cu:218-238 form)
//              chi^2 kinds   f = A1*(A0-K)^2      R = log w_s - acc * rscale_s (emc_cuda.cu:160-216, 270-275)
//              scaled chi^2  f = A1*(A0-K/phi)^2  with 1/phi[s,i] per output element (emc_cuda.cu:164-165,
//                                                 182-183, 201-205 with scaling != 1)
//   M-step : U[s,j] = sum_i P[s,i] KT[j,i]                                    (emc_atomic.cu:88-132)
//
// Both are "NT" products (reduction index contiguous in both operands).  128x128 macro tile,
// 256 threads, 8x8 register tile per thread, BK = 16, operands staged transposed in LDS so each
// thread reads float4 rows of A and B per k; global loads for tile k+1 are issued before the
// FMAs of tile k.
#include <algorithm>

#include "../emc_internal.h"

namespace emc {
namespace {

constexpr int BM = 128, BN = 128, BK = 16, LDP = BM + 4;
typedef float f32x4 __attribute__((ext_vector_type(4)));

enum { EPI_POISSON = 0, EPI_CHI2 = 1, EPI_STORE = 2, EPI_CHI2S = 3 };

template <int NA, int EPI>
__global__ __launch_bounds__(256) void k_gemm_f32(const float *__restrict__ A0, const float *__restrict__ A1,
                                                  int lda, const float *__restrict__ B, int ldb, int Kd,
                                                  float *__restrict__ C, int64_t ldc, int m_valid, int n_valid,
                                                  const float *__restrict__ bias, const float *__restrict__ rscale,
                                                  const float *__restrict__ iphi = nullptr)
{
    __shared__ float As[NA][BK][LDP];
    __shared__ float Bs[BK][LDP];
    const int tid = threadIdx.x;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int tx = tid & 15, ty = tid >> 4;

    // global -> register staging: 2 float4 per operand per thread (row = idx/4, k quad = idx%4)
    f32x4 ra[NA][2], rb[2];
    auto gload = [&](int k0) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int idx = tid + 256 * h, row = idx >> 2, kq = (idx & 3) * 4;
            ra[0][h] = *reinterpret_cast<const f32x4 *>(A0 + (size_t)(m0 + row) * lda + k0 + kq);
            if (NA == 2) ra[NA - 1][h] = *reinterpret_cast<const f32x4 *>(A1 + (size_t)(m0 + row) * lda + k0 + kq);
            rb[h] = *reinterpret_cast<const f32x4 *>(B + (size_t)(n0 + row) * ldb + k0 + kq);
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int idx = tid + 256 * h, row = idx >> 2, kq = (idx & 3) * 4;
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                As[a][kq + 0][row] = ra[a][h].x;
                As[a][kq + 1][row] = ra[a][h].y;
                As[a][kq + 2][row] = ra[a][h].z;
                As[a][kq + 3][row] = ra[a][h].w;
            }
            Bs[kq + 0][row] = rb[h].x;
            Bs[kq + 1][row] = rb[h].y;
            Bs[kq + 2][row] = rb[h].z;
            Bs[kq + 3][row] = rb[h].w;
        }
    };

    float acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = 0.0f;
    // scaled chi^2: this thread's 8 x 8 block of 1/phi (rows and columns as in the epilogue)
    float ip[EPI == EPI_CHI2S ? 8 : 1][EPI == EPI_CHI2S ? 8 : 1];
    if (EPI == EPI_CHI2S) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int m = m0 + (i < 4 ? ty * 4 + i : 64 + ty * 4 + (i - 4));
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int n = n0 + (j < 4 ? tx * 4 + j : 64 + tx * 4 + (j - 4));
                ip[i][j] = (m < m_valid && n < n_valid) ? iphi[(size_t)m * ldc + n] : 1.0f;
            }
        }
    }

    const int nk = Kd / BK;
    gload(0);
    sstore();
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        if (kt + 1 < nk) gload((kt + 1) * BK);
#pragma unroll 4
        for (int k = 0; k < BK; ++k) {
            float a0[8], a1[8], b[8];
            const f32x4 x0 = *reinterpret_cast<const f32x4 *>(&As[0][k][ty * 4]);
            const f32x4 x1 = *reinterpret_cast<const f32x4 *>(&As[0][k][64 + ty * 4]);
            a0[0] = x0.x; a0[1] = x0.y; a0[2] = x0.z; a0[3] = x0.w;
            a0[4] = x1.x; a0[5] = x1.y; a0[6] = x1.z; a0[7] = x1.w;
            if (NA == 2) {
                const f32x4 y0 = *reinterpret_cast<const f32x4 *>(&As[NA - 1][k][ty * 4]);
                const f32x4 y1 = *reinterpret_cast<const f32x4 *>(&As[NA - 1][k][64 + ty * 4]);
                a1[0] = y0.x; a1[1] = y0.y; a1[2] = y0.z; a1[3] = y0.w;
                a1[4] = y1.x; a1[5] = y1.y; a1[6] = y1.z; a1[7] = y1.w;
            }
            const f32x4 z0 = *reinterpret_cast<const f32x4 *>(&Bs[k][tx * 4]);
            const f32x4 z1 = *reinterpret_cast<const f32x4 *>(&Bs[k][64 + tx * 4]);
            b[0] = z0.x; b[1] = z0.y; b[2] = z0.z; b[3] = z0.w;
            b[4] = z1.x; b[5] = z1.y; b[6] = z1.z; b[7] = z1.w;
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    if (NA == 1) {
                        acc[i][j] = __builtin_fmaf(a0[i], b[j], acc[i][j]);
                    } else if (EPI == EPI_CHI2S) {
                        const float d = __builtin_fmaf(-b[j], ip[i][j], a0[i]);
                        acc[i][j] = __builtin_fmaf(a1[i] * d, d, acc[i][j]);
                    } else {
                        const float d = a0[i] - b[j];
                        acc[i][j] = __builtin_fmaf(a1[i] * d, d, acc[i][j]);
                    }
                }
        }
        __syncthreads();
        if (kt + 1 < nk) {
            sstore();
            __syncthreads();
        }
    }

    // epilogue
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int m = m0 + (i < 4 ? ty * 4 + i : 64 + ty * 4 + (i - 4));
        if (m >= m_valid) continue;
        float add = 0.0f, mul = 1.0f;
        if (EPI == EPI_POISSON) add = bias[m];
        if (EPI == EPI_CHI2 || EPI == EPI_CHI2S) { add = bias[m]; mul = -rscale[m]; }
#pragma unroll
        for (int jh = 0; jh < 2; ++jh) {
            const int n = n0 + jh * 64 + tx * 4;
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = __builtin_fmaf(acc[i][jh * 4 + j], mul, add);
            float *dst = C + (size_t)m * ldc + n;
            if (n + 3 < n_valid) {
                *reinterpret_cast<f32x4 *>(dst) = f32x4{v[0], v[1], v[2], v[3]};
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (n + j < n_valid) dst[j] = v[j];
            }
        }
    }
}

}  // namespace

// Logits of rotations [s0, s0+cs): A rows are the chunk buffer (chunkp rows), output rows go to
// R[s0 + r].  bias / rscale are indexed by the chunk-local row through pointer offsets.
hipError_t launch_logl_f32(const emc_ctx *c, int s0, int cs, hipStream_t st)
{
    const dim3 grid(c->Ip / BN, c->chunkp / BM), blk(256);
    float *Cp = c->R + (size_t)s0 * c->Ip;
    if (c->opt.diff == EMC_DIFF_TRUE_POISSON)
        hipLaunchKernelGGL((k_gemm_f32<1, EPI_POISSON>), grid, blk, 0, st, c->A0, nullptr, c->Pp, c->K32, c->Pp,
                           c->Pp, Cp, (int64_t)c->Ip, cs, c->I, c->bias + s0, nullptr);
    else if (c->opt.scaling)
        hipLaunchKernelGGL((k_gemm_f32<2, EPI_CHI2S>), grid, blk, 0, st, c->A0, c->A1, c->Pp, c->K32, c->Pp,
                           c->Pp, Cp, (int64_t)c->Ip, cs, c->I, c->bias + s0, c->rscale + s0,
                           c->iphi + (size_t)s0 * c->Ip);
    else
        hipLaunchKernelGGL((k_gemm_f32<2, EPI_CHI2>), grid, blk, 0, st, c->A0, c->A1, c->Pp, c->K32, c->Pp,
                           c->Pp, Cp, (int64_t)c->Ip, cs, c->I, c->bias + s0, c->rscale + s0);
    return hipGetLastError();
}

// U rows for rotations [s0, s0+cs): A = responsibilities (FP32, in R), B = counts^T.
hipError_t launch_mstep_f32(const emc_ctx *c, int s0, int cs, hipStream_t st)
{
    const dim3 grid(c->Pp / BN, c->chunkp / BM), blk(256);
    (void)cs;
    const float *A = (c->opt.scaling ? c->Pphi : c->R) + (size_t)s0 * c->Ip;  // P / phi or P
    hipLaunchKernelGGL((k_gemm_f32<1, EPI_STORE>), grid, blk, 0, st, A, nullptr, c->Ip, c->KT32, c->Ip, c->Ip, c->U,
                       (int64_t)c->Pp, c->chunkp, c->Pp, nullptr, nullptr);
    return hipGetLastError();
}

namespace {

// Operands of the scaling sums from the chunk's slices (A0 = W on valid pixels, else 0).
__global__ void k_scale_operands(const float *__restrict__ A0, int64_t n, float *__restrict__ Am,
                                 float *__restrict__ AWm)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const float w = A0[t];
        const bool in = w > 1.0e-10f;  // emc_cuda.cu:444
        Am[t] = in ? 1.0f : 0.0f;
        AWm[t] = in ? w : 0.0f;
    }
}

// 1/phi = (sum K W) / (sum K^2); rows [0, cs) of the chunk, patterns [0, I).  A pattern with no
// photon on the slice's support (either sum 0) keeps phi = 1.
__global__ void k_iphi(float *__restrict__ sumK2, const float *__restrict__ sumKW, int cs, int I, int64_t ld)
{
    const int64_t n = (int64_t)cs * ld;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        if ((int)(t % ld) >= I) continue;
        const float a = sumK2[t], b = sumKW[t];
        sumK2[t] = (a > 0.0f && b > 0.0f) ? b / a : 1.0f;
    }
}

__global__ void k_pscale(const float *__restrict__ P, const float *__restrict__ iphi, int64_t n,
                         float *__restrict__ Pphi)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        Pphi[t] = P[t] * iphi[t];
}

__global__ void k_square(const float *__restrict__ K, int64_t n, float *__restrict__ K2)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        K2[t] = K[t] * K[t];
}

unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 8192); }

}  // namespace

hipError_t launch_scaling_f32(const emc_ctx *c, int s0, int cs, hipStream_t st)
{
    const int64_t na = (int64_t)c->chunkp * c->Pp;
    hipLaunchKernelGGL(k_scale_operands, dim3(grid_for(na)), dim3(256), 0, st, c->A0, na, c->Am, c->AWm);
    const dim3 grid(c->Ip / BN, c->chunkp / BM), blk(256);
    float *sK2 = c->iphi + (size_t)s0 * c->Ip, *sKW = c->R + (size_t)s0 * c->Ip;  // R rows are free until the logits
    hipLaunchKernelGGL((k_gemm_f32<1, EPI_STORE>), grid, blk, 0, st, c->Am, nullptr, c->Pp, c->K2, c->Pp, c->Pp, sK2,
                       (int64_t)c->Ip, cs, c->Ip, nullptr, nullptr);
    hipLaunchKernelGGL((k_gemm_f32<1, EPI_STORE>), grid, blk, 0, st, c->AWm, nullptr, c->Pp, c->K32, c->Pp, c->Pp, sKW,
                       (int64_t)c->Ip, cs, c->Ip, nullptr, nullptr);
    const int64_t no = (int64_t)cs * c->Ip;
    hipLaunchKernelGGL(k_iphi, dim3(grid_for(no)), dim3(256), 0, st, sK2, sKW, cs, c->I, (int64_t)c->Ip);
    return hipGetLastError();
}

hipError_t launch_pscale_f32(const emc_ctx *c, hipStream_t st)
{
    const int64_t n = (int64_t)c->Sp * c->Ip;
    hipLaunchKernelGGL(k_pscale, dim3(grid_for(n)), dim3(256), 0, st, c->R, c->iphi, n, c->Pphi);
    return hipGetLastError();
}

hipError_t launch_square_counts(const emc_ctx *c, hipStream_t st)
{
    const int64_t n = (int64_t)c->Ip * c->Pp;
    hipLaunchKernelGGL(k_square, dim3(grid_for(n)), dim3(256), 0, st, c->K32, n, c->K2);
    return hipGetLastError();
}

}  // namespace emc
