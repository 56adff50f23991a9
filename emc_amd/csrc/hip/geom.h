// Device-side geometry shared by slicing, insertion and the fused M-step scatter.
//
// The operation sequence is part of the numerical contract: rotation matrices, rotated
// coordinates, nearest-voxel indices and trilinear fractions are computed in FP32 with explicit
// fmaf/products and no contraction, exactly as the host generator and the CPU oracle do, so
// every discrete decision (voxel index, in/out of bounds) is bit-identical across them.
#pragma once
#include <hip/hip_runtime.h>

namespace emc {

// Rotation matrix of a unit quaternion (emc_cuda.cu:113-121), row-major m[9].
__device__ __forceinline__ void rotmat(float q0, float q1, float q2, float q3, float *m)
{
#pragma clang fp contract(off)
    m[0] = __builtin_fmaf(q0, q0, __builtin_fmaf(q1, q1, __builtin_fmaf(-q2, q2, -(q3 * q3))));
    m[1] = 2.0f * __builtin_fmaf(q1, q2, -(q0 * q3));
    m[2] = 2.0f * __builtin_fmaf(q1, q3, q0 * q2);
    m[3] = 2.0f * __builtin_fmaf(q1, q2, q0 * q3);
    m[4] = __builtin_fmaf(q0, q0, __builtin_fmaf(-q1, q1, __builtin_fmaf(q2, q2, -(q3 * q3))));
    m[5] = 2.0f * __builtin_fmaf(q2, q3, -(q0 * q1));
    m[6] = 2.0f * __builtin_fmaf(q1, q3, -(q0 * q2));
    m[7] = 2.0f * __builtin_fmaf(q2, q3, q0 * q1);
    m[8] = __builtin_fmaf(q0, q0, __builtin_fmaf(-q1, q1, __builtin_fmaf(-q2, q2, q3 * q3)));
}

// u = (L/2 - 0.5) + R.(X,Y,Z): model-grid coordinate of a detector point (emc_cuda.cu:123-131)
__device__ __forceinline__ void rot_point(const float *m, float X, float Y, float Z, float c,
                                          float &ux, float &uy, float &uz)
{
#pragma clang fp contract(off)
    ux = c + __builtin_fmaf(m[0], X, __builtin_fmaf(m[1], Y, m[2] * Z));
    uy = c + __builtin_fmaf(m[3], X, __builtin_fmaf(m[4], Y, m[5] * Z));
    uz = c + __builtin_fmaf(m[6], X, __builtin_fmaf(m[7], Y, m[8] * Z));
}

// Nearest voxel, lroundf semantics (round half away from zero).
__device__ __forceinline__ int nearest(float u) { return (int)__builtin_roundf(u); }

// Trilinear cell of one axis: valid iff 0 <= u <= L-1; base clamped to L-2, fraction u - base.
__device__ __forceinline__ bool tri_axis(float u, int L, int &b, float &f)
{
#pragma clang fp contract(off)
    if (!(u >= 0.0f && u <= (float)(L - 1))) return false;
    int i0 = (int)__builtin_floorf(u);
    i0 = i0 > L - 2 ? L - 2 : i0;
    b = i0;
    f = u - (float)i0;
    return true;
}

__device__ __forceinline__ float model_c(int L) { return 0.5f * (float)L - 0.5f; }

}  // namespace emc
