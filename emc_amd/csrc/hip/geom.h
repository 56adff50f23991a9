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

// Slice value of rotation matrix m at detector point (X,Y,Z); -1 when invalid.
template <int INTERP>
__device__ __forceinline__ float sample(const float *__restrict__ model, int L, const float *m,
                                        float X, float Y, float Z)
{
    float ux, uy, uz;
    rot_point(m, X, Y, Z, model_c(L), ux, uy, uz);
    if (INTERP == 0) {
        // nearest: lroundf and the exclusive lower bound of emc_cuda.cu:129-138
        const int vx = nearest(ux), vy = nearest(uy), vz = nearest(uz);
        if (vx > 0 && vx < L && vy > 0 && vy < L && vz > 0 && vz < L)
            return model[((size_t)vz * L + vy) * L + vx];
        return -1.0f;
    } else {
        int bx, by, bz;
        float fx, fy, fz;
        if (!tri_axis(ux, L, bx, fx) || !tri_axis(uy, L, by, fy) || !tri_axis(uz, L, bz, fz))
            return -1.0f;
        const float *b = model + ((size_t)bz * L + by) * L + bx;
        const size_t LL = (size_t)L * L;
        const float v000 = b[0], v001 = b[1], v010 = b[L], v011 = b[L + 1];
        const float v100 = b[LL], v101 = b[LL + 1], v110 = b[LL + L], v111 = b[LL + L + 1];
        const float mn = fminf(fminf(fminf(v000, v001), fminf(v010, v011)),
                               fminf(fminf(v100, v101), fminf(v110, v111)));
        if (mn < 0.0f) return -1.0f;  // a missing voxel among the corners
        const float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
        const float c00 = v000 * gx + v001 * fx, c01 = v010 * gx + v011 * fx;
        const float c10 = v100 * gx + v101 * fx, c11 = v110 * gx + v111 * fx;
        const float c0 = c00 * gy + c01 * fy, c1 = c10 * gy + c11 * fy;
        return c0 * gz + c1 * fz;
    }
}

// Trilinear sample from the quad-packed model (modelq[v] = model at v, v + 1, v + L, v + L + 1; 0 past
// an edge): the 8 corners in two 16-byte loads from two cache lines instead of 8 loads from four.  The
// same corners, test and arithmetic as sample<EMC_INTERP_TRILINEAR>, so the same value bit for bit.
__device__ __forceinline__ float sample_quad(const float4 *__restrict__ modelq, int L, const float *m, float X,
                                             float Y, float Z)
{
    float ux, uy, uz;
    rot_point(m, X, Y, Z, model_c(L), ux, uy, uz);
    int bx, by, bz;
    float fx, fy, fz;
    if (!tri_axis(ux, L, bx, fx) || !tri_axis(uy, L, by, fy) || !tri_axis(uz, L, bz, fz)) return -1.0f;
    const size_t base = ((size_t)bz * L + by) * L + bx;
    const float4 q0 = modelq[base], q1 = modelq[base + (size_t)L * L];
    const float v000 = q0.x, v001 = q0.y, v010 = q0.z, v011 = q0.w;
    const float v100 = q1.x, v101 = q1.y, v110 = q1.z, v111 = q1.w;
    const float mn = fminf(fminf(fminf(v000, v001), fminf(v010, v011)),
                           fminf(fminf(v100, v101), fminf(v110, v111)));
    if (mn < 0.0f) return -1.0f;  // a missing voxel among the corners
    const float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
    const float c00 = v000 * gx + v001 * fx, c01 = v010 * gx + v011 * fx;
    const float c10 = v100 * gx + v101 * fx, c11 = v110 * gx + v111 * fx;
    const float c0 = c00 * gy + c01 * fy, c1 = c10 * gy + c11 * fy;
    return c0 * gz + c1 * fz;
}

__device__ __forceinline__ void load_rot(const float *__restrict__ rotm, int s, float *m)
{
    const float4 a = *reinterpret_cast<const float4 *>(rotm + 12 * (size_t)s);
    const float4 b = *reinterpret_cast<const float4 *>(rotm + 12 * (size_t)s + 4);
    const float c = rotm[12 * (size_t)s + 8];
    m[0] = a.x; m[1] = a.y; m[2] = a.z; m[3] = a.w;
    m[4] = b.x; m[5] = b.y; m[6] = b.z; m[7] = b.w; m[8] = c;
}

}  // namespace emc
