// One row of the sparse M-step sum: out[j] = sum_n p_n K[pat(n), j] over the non-zero entries n of
// a row of responsibilities stored as BF16 hi + lo (p_n = hi + lo).  Device code shared by the
// dense-layout and the per-cell (grouped) M-steps; one 256-thread workgroup per row.
//
// The row is scanned 8 entries per thread and step (16-byte loads); the non-zero entries are
// compacted into LDS in ascending order (wave scans and a block prefix sum keep the order, so the
// result is deterministic) and, when the list is full or the row ends, their count rows are added
// to registers, 16 pixels per thread and pass.  Counts as bytes (KT 0), BF16 (KT 1) or BF16
// hi + lo (KT 2).  FP32 accumulation.
#pragma once

#include "../emc_internal.h"

namespace emc {

namespace sparse_row_detail {
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int kCap = 4096;                  // LDS list entries per batch
constexpr int kStep = 256 * 8;              // entries scanned per step
constexpr int kPass = kSparseMaxPp / 4096;  // 16-pixel chunks per thread

__device__ __forceinline__ float half16(const u32x4 &v, int k) { return bf2f((bf16_t)(v[k >> 1] >> (16 * (k & 1)))); }
}  // namespace sparse_row_detail

// n entries of the row; pat == nullptr: entry n is pattern n.  Pp is a multiple of 16 and at most
// kSparseMaxPp; the row holds a multiple of 8 readable entries (padding entries are 0).
template <int KT>
__device__ __forceinline__ void sparse_row(const bf16_t *__restrict__ ph, const bf16_t *__restrict__ pl, int n,
                                           const int32_t *__restrict__ pat, const void *__restrict__ Kv,
                                           const bf16_t *__restrict__ Klo, int Pp, float *__restrict__ out)
{
    using namespace sparse_row_detail;
    __shared__ int li[kCap];
    __shared__ float lp[kCap];
    __shared__ int wsum[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nch = Pp / 16;
    float acc[kPass][16];
#pragma unroll
    for (int q = 0; q < kPass; ++q)
#pragma unroll
        for (int k = 0; k < 16; ++k) acc[q][k] = 0.0f;

    auto accumulate = [&](int cnt) {
#pragma unroll
        for (int q = 0; q < kPass; ++q) {
            const int pc = tid + q * 256;
            if (pc >= nch) break;
            for (int e = 0; e < cnt; ++e) {
                const size_t row = (size_t)li[e] * Pp + 16 * (size_t)pc;
                const float p = lp[e];
                if (KT == 0) {
                    const u32x4 v = *reinterpret_cast<const u32x4 *>(reinterpret_cast<const uint8_t *>(Kv) + row);
#pragma unroll
                    for (int k = 0; k < 16; ++k) acc[q][k] += p * (float)((v[k >> 2] >> (8 * (k & 3))) & 0xffu);
                } else {
                    const bf16_t *kr = reinterpret_cast<const bf16_t *>(Kv) + row;
                    const u32x4 a = *reinterpret_cast<const u32x4 *>(kr), b = *reinterpret_cast<const u32x4 *>(kr + 8);
                    u32x4 al = {0u, 0u, 0u, 0u}, bl = {0u, 0u, 0u, 0u};
                    if (KT == 2) {
                        al = *reinterpret_cast<const u32x4 *>(Klo + row);
                        bl = *reinterpret_cast<const u32x4 *>(Klo + row + 8);
                    }
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        float k0 = half16(a, k), k1 = half16(b, k);
                        if (KT == 2) {
                            k0 += half16(al, k);
                            k1 += half16(bl, k);
                        }
                        acc[q][k] += p * k0;
                        acc[q][8 + k] += p * k1;
                    }
                }
            }
        }
    };

    int m = 0;  // entries in the list (the same in every thread)
    for (int i0 = 0; i0 < n; i0 += kStep) {
        const int ib = i0 + tid * 8;
        u32x4 h = {0u, 0u, 0u, 0u}, l = {0u, 0u, 0u, 0u};
        if (ib < n) {
            h = *reinterpret_cast<const u32x4 *>(ph + ib);
            l = *reinterpret_cast<const u32x4 *>(pl + ib);
        }
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t bits = ((h[k >> 1] | l[k >> 1]) >> (16 * (k & 1))) & 0xffffu;
            cnt += (bits != 0u && ib + k < n) ? 1 : 0;
        }
        int incl = cnt;  // inclusive scan over the wave
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        int woff = 0;
        for (int w = 0; w < wave; ++w) woff += wsum[w];
        const int step_total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (m + step_total > kCap) {  // block-uniform
            accumulate(m);
            __syncthreads();
            m = 0;
        }
        int pos = m + woff + incl - cnt;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint16_t hb = (uint16_t)(h[k >> 1] >> (16 * (k & 1))), lb = (uint16_t)(l[k >> 1] >> (16 * (k & 1)));
            if ((hb | lb) != 0 && ib + k < n) {
                li[pos] = pat ? pat[ib + k] : ib + k;
                lp[pos] = bf2f(hb) + bf2f(lb);
                ++pos;
            }
        }
        m += step_total;
        __syncthreads();
    }
    accumulate(m);

#pragma unroll
    for (int q = 0; q < kPass; ++q) {
        const int pc = tid + q * 256;
        if (pc >= nch) break;
#pragma unroll
        for (int k = 0; k < 16; k += 4)
            *reinterpret_cast<float4 *>(out + 16 * pc + k) = make_float4(acc[q][k], acc[q][k + 1], acc[q][k + 2], acc[q][k + 3]);
    }
}

}  // namespace emc
