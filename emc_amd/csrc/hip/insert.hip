// Maximize, scatter half: insert weighted slice sums into the model numerator and weight with
// float atomics (emc_atomic.cu:33-85, 183-200).
//
// U[r, j] = sum_i P[s, i] K[i, j] for rotation s = s0 + r (the reference's update_slices_kernel
// output times T_s).  Slices with T_s <= 1e-10 are skipped (emc_atomic.cu:190); pixels with
// U < 0 are skipped (emc_atomic.cu:56).  Nearest uses 0 <= v < L (emc_atomic.cu:67-69);
// trilinear splats to the 8 corners with the slicing weights (the adjoint of trilinear slicing).
// The model is L2-resident at 64^3, so the atomics resolve in L2.
#include "../emc_internal.h"
#include "geom.h"

namespace emc {
namespace {

template <int INTERP>
__global__ __launch_bounds__(256) void k_insert(const float *__restrict__ U, int Pp, int s0, int cs,
                                                const double *__restrict__ T, const float *__restrict__ rotm,
                                                const float *__restrict__ axyz, int Pa, int L,
                                                float *__restrict__ Mnum, float *__restrict__ Wt)
{
    const int r = blockIdx.x;
    if (r >= cs) return;
    const int s = s0 + r;
    const double Td = T[s];
    if (!(Td > 1e-10)) return;
    const float Ts = (float)Td;
    float m[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = rotm[12 * (size_t)s + k];
    const float c = model_c(L);
    for (int j = threadIdx.x; j < Pa; j += 256) {
        const float u = U[(size_t)r * Pp + j];
        if (!(u >= 0.0f)) continue;
        float ux, uy, uz;
        rot_point(m, axyz[j], axyz[Pp + j], axyz[2 * Pp + j], c, ux, uy, uz);
        if (INTERP == EMC_INTERP_NEAREST) {
            const int vx = nearest(ux), vy = nearest(uy), vz = nearest(uz);
            if (vx >= 0 && vx < L && vy >= 0 && vy < L && vz >= 0 && vz < L) {
                const size_t v = ((size_t)vz * L + vy) * L + vx;
                atomicAdd(Mnum + v, u);
                atomicAdd(Wt + v, Ts);
            }
        } else {
            int bx, by, bz;
            float fx, fy, fz;
            if (!tri_axis(ux, L, bx, fx) || !tri_axis(uy, L, by, fy) || !tri_axis(uz, L, bz, fz)) continue;
            const float g[2][3] = {{1.0f - fx, 1.0f - fy, 1.0f - fz}, {fx, fy, fz}};
            const size_t base = ((size_t)bz * L + by) * L + bx;
#pragma unroll
            for (int dz = 0; dz < 2; ++dz)
#pragma unroll
                for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                    for (int dx = 0; dx < 2; ++dx) {
                        const float w = g[dx][0] * g[dy][1] * g[dz][2];
                        const size_t v = base + ((size_t)dz * L + dy) * L + dx;
                        atomicAdd(Mnum + v, w * u);
                        atomicAdd(Wt + v, w * Ts);
                    }
        }
    }
}

}  // namespace

hipError_t launch_insert(const emc_ctx *c, int s0, int cs, hipStream_t st)
{
    if (cs <= 0) return hipSuccess;
    if (c->opt.interp == EMC_INTERP_NEAREST)
        hipLaunchKernelGGL(k_insert<EMC_INTERP_NEAREST>, dim3(cs), dim3(256), 0, st, c->U, c->Pp, s0, cs, c->T,
                           c->rotm, c->axyz, c->Pa, c->L, c->Mnum, c->Wt);
    else
        hipLaunchKernelGGL(k_insert<EMC_INTERP_TRILINEAR>, dim3(cs), dim3(256), 0, st, c->U, c->Pp, s0, cs, c->T,
                           c->rotm, c->axyz, c->Pa, c->L, c->Mnum, c->Wt);
    return hipGetLastError();
}

}  // namespace emc
