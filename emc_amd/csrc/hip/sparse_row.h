// One row of the sparse M-step sum: out[j] = sum_n p_n K[pat(n), j] over the non-zero entries n of
// a row of responsibilities stored as BF16 hi + lo (p_n = hi + lo).  Device code shared by the
// dense-layout and the per-cell (grouped) M-steps; one 256-thread workgroup per row.
//
// The row is scanned 32 entries per thread and step (two 16-byte loads of hi per thread in flight;
// lo only beside a non-zero hi, which is exact for the split lo = bf16(p - hi)).  A step with no
// non-zero hi in the whole workgroup ends after one barrier; otherwise its non-zero entries are
// compacted into an LDS list in ascending order (wave scans and a block prefix sum keep the order,
// so the result is deterministic), in as many batches as the list needs, and each full list (and
// the last) has its count rows added to registers, 16 pixels per thread and pass.  Counts as
// bytes (KT 0), BF16 (KT 1) or BF16 hi + lo (KT 2); NPASS 16-pixel chunks per thread (Pp <=
// 4096 * NPASS).  FP32 accumulation.
#pragma once

#include "../emc_internal.h"

namespace emc {

namespace sparse_row_detail {
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int kCap = 2048;        // LDS list entries per batch
constexpr int kV = 4;             // 16-byte loads of hi per thread and step
constexpr int kPer = 8 * kV;      // consecutive entries per thread and step
constexpr int kStep = 256 * kPer; // entries scanned per step
static_assert(256 % kPer == 0, "a thread's entries lie in one 256-entry segment");

__device__ __forceinline__ float half16(const u32x4 &v, int k) { return bf2f((bf16_t)(v[k >> 1] >> (16 * (k & 1)))); }
__device__ __forceinline__ bool nonzero(const u32x4 &v) { return (v[0] | v[1] | v[2] | v[3]) != 0u; }
}  // namespace sparse_row_detail

// n entries of the row; pat == nullptr: entry n is pattern n.  Pp is a multiple of 16 and at most
// 4096 * NPASS; the row holds a multiple of 8 readable entries (padding entries are 0) and starts
// 16-byte aligned.  seg (or nullptr): one flag per 256 entries of the row, 0 where those entries are
// known to be zero (their loads are skipped).
template <int KT, int NPASS>
__device__ __forceinline__ void sparse_row(const bf16_t *__restrict__ ph, const bf16_t *__restrict__ pl, int n,
                                           const int32_t *__restrict__ pat, const void *__restrict__ Kv,
                                           const bf16_t *__restrict__ Klo, int Pp, float *__restrict__ out,
                                           const uint8_t *__restrict__ seg = nullptr)
{
    using namespace sparse_row_detail;
    __shared__ int li[kCap];
    __shared__ float lp[kCap];
    __shared__ int wsum[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nch = Pp / 16;
    float acc[NPASS][16];
#pragma unroll
    for (int q = 0; q < NPASS; ++q)
#pragma unroll
        for (int k = 0; k < 16; ++k) acc[q][k] = 0.0f;

    auto accumulate = [&](int cnt) {
#pragma unroll
        for (int q = 0; q < NPASS; ++q) {
            const int pc = tid + q * 256;
            if (pc >= nch) break;
            // unrolled so that the count-row loads of several entries are in flight together (the
            // sums still add the entries in list order)
#pragma unroll 4
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
        const int ib = i0 + tid * kPer;  // kPer divides 256: the thread's entries share one flag
        const bool live = ib < n && (!seg || seg[ib >> 8] != 0);
        u32x4 h[kV], l[kV];
        bool any = false;
#pragma unroll
        for (int v = 0; v < kV; ++v) {
            h[v] = u32x4{0u, 0u, 0u, 0u};
            l[v] = u32x4{0u, 0u, 0u, 0u};
            if (live && ib + 8 * v < n) h[v] = *reinterpret_cast<const u32x4 *>(ph + ib + 8 * v);
            any |= nonzero(h[v]);
        }
        if (!__syncthreads_or(any)) continue;
        int cnt = 0;
#pragma unroll
        for (int v = 0; v < kV; ++v) {
            if (nonzero(h[v])) l[v] = *reinterpret_cast<const u32x4 *>(pl + ib + 8 * v);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t bits = ((h[v][k >> 1] | l[v][k >> 1]) >> (16 * (k & 1))) & 0xffffu;
                cnt += (bits != 0u && ib + 8 * v + k < n) ? 1 : 0;
            }
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
        const int rank0 = woff + incl - cnt;  // rank of this thread's first entry within the step
        // the step's entries with rank in [base, base + take) fill the list from position m
        for (int base = 0;;) {  // block-uniform loop
            const int take = min(kCap - m, step_total - base);
            int rank = rank0;
#pragma unroll
            for (int v = 0; v < kV; ++v)
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const uint16_t hb = (uint16_t)(h[v][k >> 1] >> (16 * (k & 1)));
                    const uint16_t lb = (uint16_t)(l[v][k >> 1] >> (16 * (k & 1)));
                    if ((hb | lb) != 0 && ib + 8 * v + k < n) {
                        const int q = rank - base;
                        if (q >= 0 && q < take) {
                            li[m + q] = pat ? pat[ib + 8 * v + k] : ib + 8 * v + k;
                            lp[m + q] = bf2f(hb) + bf2f(lb);
                        }
                        ++rank;
                    }
                }
            m += take;
            base += take;
            __syncthreads();
            if (base >= step_total) break;
            accumulate(m);  // the list is full
            __syncthreads();
            m = 0;
        }
    }
    accumulate(m);

#pragma unroll
    for (int q = 0; q < NPASS; ++q) {
        const int pc = tid + q * 256;
        if (pc >= nch) break;
#pragma unroll
        for (int k = 0; k < 16; k += 4)
            *reinterpret_cast<float4 *>(out + 16 * pc + k) = make_float4(acc[q][k], acc[q][k + 1], acc[q][k + 2], acc[q][k + 3]);
    }
}

}  // namespace emc
