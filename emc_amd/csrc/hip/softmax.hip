// Responsibilities: per-pattern log-sum-exp over rotations and P = exp(R - lse).
//
// Restates cuda_max_vector + cuda_respons_max_expf + the host log/add + cuda_norm_respons_sumexpf
// (emc_cuda.cu:1041-1056, 1175-1217, 1253-1272; main.cpp:367-392) as three streaming passes over
// the logits R[S][I] (row = rotation, so a pattern is a column):
//   1. partial: each thread owns one pattern column over a band of rotation rows and keeps an
//      online (max, sum of exp) pair; exp is FP32, the sum FP64.  Reads are coalesced rows.
//   2. final:   combine the bands per pattern in FP64 -> lse'_i; per-block partial sums of lse'.
//   3. apply:   P = exp(R - lse'_i) written either in place (FP32 path) or as a BF16 hi+lo pair
//      (MFMA path), plus per-band row sums of P (T_s partials), reduced in FP64, and (MFMA path)
//      the count of non-zero P, which picks the dense or the sparse form of the M-step.
// The reference's -1e10 floor on the max and on R - max is kept.
#include <algorithm>

#include "../emc_internal.h"

namespace emc {
namespace {

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(256) void k_sm_partial(const float *__restrict__ R, int64_t ldr, int S, int I,
                                                    int rows_per_band, float *__restrict__ pm,
                                                    double *__restrict__ ps)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int band = blockIdx.y;
    if (i >= I) return;
    const int r0 = band * rows_per_band, r1 = min(S, r0 + rows_per_band);
    float m = -1.0e10f;
    double sum = 0.0;
    const float *col = R + i;
    for (int s = r0; s < r1; ++s) {
        const float v = col[(size_t)s * ldr];
        if (v > m) {
            sum = sum * (double)expf(m - v) + 1.0;
            m = v;
        } else {
            const float d = v - m;
            if (d > -1.0e10f) sum += (double)expf(d);
        }
    }
    pm[(size_t)band * I + i] = m;
    ps[(size_t)band * I + i] = sum;
}

// lse_i = m_i + log sum_b ps_b exp(pm_b - m_i) over the bands.  Eight lanes per pattern take every
// eighth band and combine over a fixed butterfly (deterministic); 32 patterns per workgroup.
constexpr int kFinalPat = 32;
__global__ __launch_bounds__(256) void k_sm_final(const float *__restrict__ pm, const double *__restrict__ ps,
                                                  int nbands, int I, double *__restrict__ lse,
                                                  double *__restrict__ blk_sum, double *__restrict__ scal)
{
    __shared__ double red[4];
    const int g = threadIdx.x & 7;
    const int i = blockIdx.x * kFinalPat + (threadIdx.x >> 3);
    float m = -1.0e10f;
    if (i < I)
        for (int b = g; b < nbands; b += 8) m = fmaxf(m, pm[(size_t)b * I + i]);
    m = fmaxf(m, __shfl_xor(m, 1, 64));
    m = fmaxf(m, __shfl_xor(m, 2, 64));
    m = fmaxf(m, __shfl_xor(m, 4, 64));
    double sum = 0.0;
    if (i < I)
        for (int b = g; b < nbands; b += 8) {
            const double sb = ps[(size_t)b * I + i];
            if (sb > 0.0) sum += sb * exp((double)pm[(size_t)b * I + i] - (double)m);
        }
    sum += __shfl_xor(sum, 1, 64);
    sum += __shfl_xor(sum, 2, 64);
    sum += __shfl_xor(sum, 4, 64);
    double l = 0.0;
    if (i < I && g == 0) {
        l = (double)m + log(sum);
        if (!isfinite(l)) scal[SC_NONFINITE] = 1.0;
        lse[i] = l;
    }
    l = wave_sum(l);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) blk_sum[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Deterministic single-workgroup sum of n doubles into *out.
__global__ __launch_bounds__(1024) void k_sum_f64(const double *__restrict__ x, int n, double *__restrict__ out)
{
    __shared__ double red[16];
    // eight partial sums per thread (eight loads in flight), added in a fixed order
    double p[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int k = threadIdx.x;
    for (; k + 7 * 1024 < n; k += 8 * 1024)
#pragma unroll
        for (int q = 0; q < 8; ++q) p[q] += x[k + q * 1024];
#pragma unroll
    for (int q = 0; q < 8; ++q)
        if (k + q * 1024 < n) p[q] += x[k + q * 1024];
    double a = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
    a = wave_sum(a);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < 16; ++k) t += red[k];
        *out = t;
    }
}

constexpr int kApplyRows = 8;    // rotation rows per workgroup (the block reads of k_sm_apply assume 8)

// Two floats to a packed BF16 pair (a in the low half), round to nearest even: gfx950's
// two-source v_cvt_pk_bf16_f32, which rounds as f2bf_dev does.  Written as asm because the compiler
// emits one single-source conversion per value and merges the halves with extra instructions.
// The s_nop is the wait state gfx950 requires between a transcendental VALU instruction (the
// v_exp_f32 that produces p) and a VALU instruction that reads its result, and between a write with
// a destination select and its reader: the compiler inserts these for the instructions it emits but
// not in front of inline asm, and without it the conversion can read a stale value when scheduled
// right after the exp.
__device__ __forceinline__ uint32_t pk_bf16(float a, float b)
{
    uint32_t r;
    asm("s_nop 0\n\tv_cvt_pk_bf16_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
constexpr int kApplyCols = 1024; // pattern columns per workgroup (4 per thread)

// OUT = 0: P in place (FP32).  OUT = 1: BF16 hi/lo pair, of P * iphi when iphi is given (the
// M-step's P/phi under intensity scaling; the slice sums T stay sums of P).
// OUT = 1 skips the store of a wave's 256-pattern row segment when every value in it is zero and
// pdirty says the segment already holds only zeros (P is mostly zeros once the posteriors are
// sharp); pdirty is then updated to whether the segment holds a non-zero value.
// pm (the logL epilogue's band maxima, [band][I], band of row s = (s / chunk) bpc + (s % chunk) / 256;
// null without them): a pattern's logits in a band whose maximum lies more than kSkipApplyNats below
// its lse are not read and count as -inf, which gives the P = 0 that they give (exp of less than -110
// is 0 in FP32).  blk (the skipping logL pass's record, [band][I]; null when every logit is in
// place): -1 = the band's logits of the pattern were not stored (they are negligible: the pass
// leaves out only bands more than 128 below a maximum of the pattern, so pm marks them too), -2 = in
// place, k >= 0 = the 256 logits of the band are the row brow0 + k of R from the pattern's 256-column
// tile on (brow0 = the band's first row).
constexpr double kSkipApplyNats = 110.0;
template <int OUT>
__global__ __launch_bounds__(256) void k_sm_apply(float *__restrict__ R, int64_t ldr, int S, int I,
                                                  const double *__restrict__ lse, bf16_t *__restrict__ Ph,
                                                  bf16_t *__restrict__ Pl, double *__restrict__ Tpart,
                                                  double *__restrict__ nnz, const float *__restrict__ iphi,
                                                  uint8_t *__restrict__ pdirty, int nseg, const float *__restrict__ pm,
                                                  const int16_t *__restrict__ blk, int chunk, int bpc)
{
    static_assert(kApplyRows == 8, "8 rows: two per wave in the row sums, 8 floats per block read");
    constexpr float kLog2e = 1.4426950408889634f;
    int nz = 0;  // OUT = 1: entries with a non-zero hi or lo term (for the choice of the M-step)
    // per-thread row sums: each is a float (a float sum of the thread's four values, or one value),
    // kept as such and widened when the row sums are formed
    __shared__ float rs[kApplyRows][256];
    const int c0 = blockIdx.x * kApplyCols + threadIdx.x * 4;
    const int r0 = blockIdx.y * kApplyRows;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int seg = (blockIdx.x * kApplyCols) / 256 + wave;  // this wave's 256 patterns
    const bool seg_live = seg * 256 < I;                      // uniform over the wave
    // lse as a float pair: x - lh is exact wherever p is not negligible (x and lse within a factor
    // of 2, Sterbenz), so exp((x - lh) - ll) carries the double-precision argument to float rounding
    // (a column past I gets lh = +inf: its p is exp2(-inf) = 0 without a select per element)
    // The thread's four patterns' lse, band maxima and block records are read with one vector load
    // each (lse always: c0 is a multiple of 4 and lse holds Ip >= I rounded up to 4 entries; the
    // [band][I] rows of pm and blk when I is a multiple of 4), element by element otherwise: the
    // short workgroups of this kernel are bound by the vector-memory data path.
    double lv[4];
    if (c0 < I) {
        const double2 a = *reinterpret_cast<const double2 *>(lse + c0), b = *reinterpret_cast<const double2 *>(lse + c0 + 2);
        lv[0] = a.x; lv[1] = a.y; lv[2] = b.x; lv[3] = b.y;
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) lv[k] = 0.0;
    }
    float lh[4], ll[4];
    bool neg[4];  // this band's logits of pattern c0 + k are negligible (not read)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double l = (c0 + k < I) ? lv[k] : 0.0;
        lh[k] = (c0 + k < I) ? (float)l : INFINITY;
        ll[k] = (c0 + k < I) ? (float)(l - (double)lh[k]) : 0.0f;
        neg[k] = false;
    }
    int code[4] = {-2, -2, -2, -2};
    int brow0 = 0, b0 = 0;
    bool one_band = false;
    if (OUT == 1 && pm) {
        const int rl = min(r0 + kApplyRows - 1, S - 1);
        b0 = (r0 / chunk) * bpc + (r0 % chunk) / 256;
        one_band = b0 == (rl / chunk) * bpc + (rl % chunk) / 256;
        if (one_band) {  // the workgroup's rows lie in one band (always, when blk is given)
            brow0 = (r0 / chunk) * chunk + ((r0 % chunk) / 256) * 256;
            float pv[4];
            int cv[4] = {-2, -2, -2, -2};
            const size_t o = (size_t)b0 * I + c0;
            if ((I & 3) == 0 && c0 < I) {  // uniform layout test; then c0 + 3 < I
                const float4 v = *reinterpret_cast<const float4 *>(pm + o);
                pv[0] = v.x; pv[1] = v.y; pv[2] = v.z; pv[3] = v.w;
                if (blk) {
                    const uint2 w = *reinterpret_cast<const uint2 *>(blk + o);
                    cv[0] = (int16_t)(w.x & 0xffffu); cv[1] = (int16_t)(w.x >> 16);
                    cv[2] = (int16_t)(w.y & 0xffffu); cv[3] = (int16_t)(w.y >> 16);
                }
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    pv[k] = c0 + k < I ? pm[o + k] : 0.0f;
                    if (blk && c0 + k < I) cv[k] = blk[o + k];
                }
            }
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (c0 + k < I) {
                    neg[k] = (double)pv[k] < lv[k] - kSkipApplyNats;
                    code[k] = cv[k];
                    neg[k] = neg[k] || code[k] == -1;
                }
        }
    }
    // segment flags (a segment past the last reads the last; rows past S read row S - 1: not used):
    // the same 8 bytes for the whole wavefront, so lane rr loads row rr's and the others take it by
    // readlane
    int flag[kApplyRows];
    const int segc = OUT == 1 ? min(seg, nseg - 1) : 0;
    {
        const int fv = OUT == 1 && lane < kApplyRows ? (int)pdirty[(size_t)min(r0 + lane, S - 1) * nseg + segc] : 0;
#pragma unroll
        for (int rr = 0; rr < kApplyRows; ++rr) flag[rr] = __builtin_amdgcn_readlane(fv, rr);
    }

    // Sparse form: when at most kSparseMax of the workgroup's patterns have a band that is not
    // negligible, one thread per such pattern computes its 8 values (the others are exactly 0), the
    // values go to LDS, and the P row segments that have to be written are assembled from there.
    constexpr int kSparseMax = 256;
    __shared__ int16_t amap[kApplyCols];               // local pattern -> its slot, -1 if negligible
    __shared__ int16_t alist[kSparseMax];              // slot -> local pattern
    __shared__ uint32_t aval[kSparseMax][kApplyRows];  // hi | lo << 16 per slot and row
    __shared__ int wcnt[4];
    __shared__ uint32_t nzf[kApplyRows];               // bit g: segment g of the row holds a non-zero
    bool sparse = false;
    if (OUT == 1 && pm && !iphi && one_band) {  // uniform over the workgroup
        bool act[4];
        int below = 0, wc = 0;
        const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            act[k] = c0 + k < I && !neg[k];
            const uint64_t bal = __ballot(act[k]);
            wc += __popcll(bal);
            below += __popcll(bal & lt);
        }
        if (lane == 0) wcnt[wave] = wc;
        if (threadIdx.x < kApplyRows) nzf[threadIdx.x] = 0u;
        __syncthreads();
        const int total = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        sparse = total <= kSparseMax;
        if (sparse) {
            int slot = below;
            for (int w = 0; w < wave; ++w) slot += wcnt[w];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                amap[4 * threadIdx.x + k] = (int16_t)(act[k] ? slot : -1);
                if (act[k]) alist[slot++] = (int16_t)(4 * threadIdx.x + k);
            }
            __syncthreads();
            float ts[kApplyRows];
#pragma unroll
            for (int rr = 0; rr < kApplyRows; ++rr) ts[rr] = 0.0f;
            if (threadIdx.x < total) {
                const int lp = alist[threadIdx.x], n = blockIdx.x * kApplyCols + lp;
                const double l = lse[n];
                const float lhn = (float)l, lln = (float)(l - (double)lhn);
                const int cd = blk ? (int)blk[(size_t)b0 * I + n] : -2;
                float x[kApplyRows];
                if (cd >= 0) {
                    const float *b = R + (size_t)(brow0 + cd) * ldr + (n & ~255) + (r0 - brow0);
                    const float4 a0 = *reinterpret_cast<const float4 *>(b), a1 = *reinterpret_cast<const float4 *>(b + 4);
                    x[0] = a0.x; x[1] = a0.y; x[2] = a0.z; x[3] = a0.w;
                    x[4] = a1.x; x[5] = a1.y; x[6] = a1.z; x[7] = a1.w;
                } else {
#pragma unroll
                    for (int rr = 0; rr < kApplyRows; ++rr) x[rr] = R[(size_t)min(r0 + rr, S - 1) * ldr + n];
                }
#pragma unroll
                for (int rr = 0; rr < kApplyRows; ++rr) {
                    const float p = r0 + rr < S ? __builtin_amdgcn_exp2f(((x[rr] - lhn) - lln) * kLog2e) : 0.0f;
                    ts[rr] = p;
                    // the dense form's rounding: hi = bf16(p), lo = bf16(p - hi), one conversion each
                    const uint32_t h = pk_bf16(p, 0.0f) & 0xffffu;
                    const uint32_t lo = pk_bf16(p - __builtin_bit_cast(float, h << 16), 0.0f) & 0xffffu;
                    aval[threadIdx.x][rr] = h | (lo << 16);
                    if (h != 0u) {
                        ++nz;
                        atomicOr(&nzf[rr], 1u << (lp >> 8));
                    }
                }
            }
#pragma unroll
            for (int rr = 0; rr < kApplyRows; ++rr) rs[rr][threadIdx.x] = ts[rr];
        }
    }
    if (!sparse) {
        // every row's logits are loaded before the first is used, so the wave waits for memory once
        // rather than once per row (rows past S read row S - 1: not used)
        float xv[kApplyRows][4];
        const bool in_place = code[0] == -2 && code[1] == -2 && code[2] == -2 && code[3] == -2;
        const bool read = !(neg[0] && neg[1] && neg[2] && neg[3]);
#pragma unroll
        for (int rr = 0; rr < kApplyRows; ++rr) {
            const size_t sc = (size_t)min(r0 + rr, S - 1);
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (c0 < I && read && in_place) v = *reinterpret_cast<const float4 *>(R + sc * ldr + c0);
            xv[rr][0] = v.x;
            xv[rr][1] = v.y;
            xv[rr][2] = v.z;
            xv[rr][3] = v.w;
        }
        if (!in_place)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (neg[k]) continue;
                if (code[k] >= 0) {  // 8 consecutive floats of the pattern's block
                    const float *b = R + (size_t)(brow0 + code[k]) * ldr + ((c0 + k) & ~255) + (r0 - brow0);
                    const float4 a0 = *reinterpret_cast<const float4 *>(b), a1 = *reinterpret_cast<const float4 *>(b + 4);
                    xv[0][k] = a0.x; xv[1][k] = a0.y; xv[2][k] = a0.z; xv[3][k] = a0.w;
                    xv[4][k] = a1.x; xv[5][k] = a1.y; xv[6][k] = a1.z; xv[7][k] = a1.w;
                } else {
#pragma unroll
                    for (int rr = 0; rr < kApplyRows; ++rr) xv[rr][k] = R[(size_t)min(r0 + rr, S - 1) * ldr + c0 + k];
                }
            }
#pragma unroll
        for (int rr = 0; rr < kApplyRows; ++rr) {
            const int s = r0 + rr;
            float acc = 0.0f;
            uint2 hv = make_uint2(0u, 0u), lv = make_uint2(0u, 0u);
            int any_nz = 0;
            if (s < S && c0 < I) {
                float *src = R + (size_t)s * ldr + c0;
                const float x[4] = {neg[0] ? -INFINITY : xv[rr][0], neg[1] ? -INFINITY : xv[rr][1],
                                    neg[2] ? -INFINITY : xv[rr][2], neg[3] ? -INFINITY : xv[rr][3]};
                // exp(d) = 2^(d log2 e) on the hardware exp2: relative error about |d| 6e-8, below the
                // 2^-16 of the BF16 pair P is stored as
                float p[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) p[k] = __builtin_amdgcn_exp2f(((x[k] - lh[k]) - ll[k]) * kLog2e);
                acc = (p[0] + p[1]) + (p[2] + p[3]);
                if (OUT == 0) {
                    *reinterpret_cast<float4 *>(src) = make_float4(p[0], p[1], p[2], p[3]);
                } else {
                    const size_t o = (size_t)s * ldr + c0;
                    float q[4] = {p[0], p[1], p[2], p[3]};
                    if (iphi) {
                        const float4 f = *reinterpret_cast<const float4 *>(iphi + o);
                        q[0] *= f.x; q[1] *= f.y; q[2] *= f.z; q[3] *= f.w;
                    }
                    // hi of a pair in one v_cvt_pk_bf16_f32; its two values back as floats by a shift
                    // and a mask; lo = bf16(q - hi).  q >= 0, so (hi | lo) != 0 exactly when hi != 0.
                    const uint32_t h01 = pk_bf16(q[0], q[1]), h23 = pk_bf16(q[2], q[3]);
                    const uint32_t l01 = pk_bf16(q[0] - __builtin_bit_cast(float, h01 << 16),
                                                 q[1] - __builtin_bit_cast(float, h01 & 0xffff0000u));
                    const uint32_t l23 = pk_bf16(q[2] - __builtin_bit_cast(float, h23 << 16),
                                                 q[3] - __builtin_bit_cast(float, h23 & 0xffff0000u));
                    const int mine = ((h01 & 0xffffu) != 0) + ((h01 >> 16) != 0) + ((h23 & 0xffffu) != 0) +
                                     ((h23 >> 16) != 0);
                    nz += mine;
                    hv = make_uint2(h01, h23);
                    lv = make_uint2(l01, l23);
                    any_nz = mine != 0;
                }
            }
            if (OUT == 1 && s < S && seg_live) {  // s and seg_live are uniform over the wave
                const bool seg_nz = __any(any_nz) != 0;
                const size_t fo = (size_t)s * nseg + seg;
                const bool dirty = flag[rr] != 0;
                if ((seg_nz || dirty) && c0 < I) {
                    const size_t o = (size_t)s * ldr + c0;
                    *reinterpret_cast<uint2 *>(Ph + o) = hv;
                    *reinterpret_cast<uint2 *>(Pl + o) = lv;
                }
                if (lane == 0 && dirty != seg_nz) pdirty[fo] = seg_nz ? 1 : 0;
            }
            rs[rr][threadIdx.x] = acc;
        }
    }
    // row sums over the workgroup's 1,024 patterns through LDS: wave w adds rows 2w and 2w + 1 (one
    // per 32-lane half, eight values a lane, then five shuffle steps), in a fixed order
    __shared__ int nzw[4];
    if (OUT == 1) {
        nz = wave_sum(nz);
        if (lane == 0) nzw[wave] = nz;
    }
    __syncthreads();
    {
        const int row = 2 * wave + (lane >> 5), part = lane & 31;
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q) t += (double)rs[row][part + 32 * q];
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
        if (part == 0 && r0 + row < S) Tpart[(size_t)blockIdx.x * S + r0 + row] = t;
    }
    if (sparse) {  // the P row segments that hold a non-zero or held one, assembled from the slots
#pragma unroll
        for (int rr = 0; rr < kApplyRows; ++rr) {
            const int s = r0 + rr;
            if (s >= S || !seg_live) continue;
            const bool seg_nz = (nzf[rr] >> wave) & 1u;
            const bool dirty = flag[rr] != 0;
            if ((seg_nz || dirty) && c0 < I) {
                uint32_t v[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int a = amap[4 * threadIdx.x + k];
                    v[k] = a >= 0 ? aval[a][rr] : 0u;
                }
                const size_t o = (size_t)s * ldr + c0;
                *reinterpret_cast<uint2 *>(Ph + o) =
                    make_uint2((v[0] & 0xffffu) | (v[1] << 16), (v[2] & 0xffffu) | (v[3] << 16));
                *reinterpret_cast<uint2 *>(Pl + o) =
                    make_uint2((v[0] >> 16) | (v[1] & 0xffff0000u), (v[2] >> 16) | (v[3] & 0xffff0000u));
            }
            if (lane == 0 && dirty != seg_nz) pdirty[(size_t)s * nseg + seg] = seg_nz ? 1 : 0;
        }
    }
    // per-workgroup non-zero count, summed by one deterministic pass afterwards
    if (OUT == 1 && threadIdx.x == 0)
        nnz[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = (double)(nzw[0] + nzw[1] + nzw[2] + nzw[3]);
}

// T_s = sum over column blocks of the partial row sums (deterministic order).
__global__ void k_T_final(const double *__restrict__ Tpart, int ncb, int S, int Sp, double *__restrict__ T)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= Sp) return;
    double t = 0.0;
    if (s < S)
        for (int b = 0; b < ncb; ++b) t += Tpart[(size_t)b * S + s];
    T[s] = t;
}

}  // namespace

// Partials layout inside c->part:  pm (float, nb*I) | ps (double, nb*I) | blk (double) | Tpart |
// non-zero counts (one per apply workgroup).
// nb = the partial pass's band count, or (MFMA logL with fused partials) one band per 256 rows.
int softmax_bands(int S, int I)
{
    return std::max(1, std::min((S + 63) / 64, 4096 / std::max(1, (I + 255) / 256) + 1));
}

int64_t softmax_part_len(int S, int I, int fused_bands)
{
    const int64_t nb = std::max(softmax_bands(S, I), fused_bands);
    const int ncb = (I + kApplyCols - 1) / kApplyCols;
    return nb * I /*pm as floats packed in doubles: over-allocate*/ + nb * I + (I + kFinalPat - 1) / kFinalPat + 1 +
           (int64_t)ncb * S + (int64_t)ncb * ((S + kApplyRows - 1) / kApplyRows) + 16;
}

__global__ void k_mark_dirty(uint8_t *__restrict__ d, int64_t n)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        d[t] = 1;
}

hipError_t launch_mark_p_dirty(const emc_ctx *c, hipStream_t st)
{
    if (!c->pdirty) return hipSuccess;
    const int64_t n = (int64_t)c->Sp * p_segments(c);
    hipLaunchKernelGGL(k_mark_dirty, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)), dim3(256), 0, st,
                       c->pdirty, n);
    return hipGetLastError();
}

hipError_t launch_sum_f64(const double *x, int n, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(1024), 0, st, x, n, out);
    return hipGetLastError();
}

hipError_t launch_softmax(const emc_ctx *c, hipStream_t st)
{
    const int S = c->S, I = c->I;
    const int ncol = (I + 255) / 256;
    const int nfin = (I + kFinalPat - 1) / kFinalPat;
    const bool fused = logl_fuses_softmax(c);  // the logL epilogue wrote the band partials
    const int nb = fused ? softmax_fused_bands(c) : softmax_bands(S, I);
    const int rows = (S + nb - 1) / nb;
    double *base = c->part;
    float *pm = reinterpret_cast<float *>(base);
    double *ps = base + (int64_t)nb * I;
    double *blk = ps + (int64_t)nb * I;
    double *Tpart = blk + nfin + 1;
    if (!fused)
        hipLaunchKernelGGL(k_sm_partial, dim3(ncol, nb), dim3(256), 0, st, c->R, (int64_t)c->Ip, S, I, rows, pm, ps);
    hipLaunchKernelGGL(k_sm_final, dim3(nfin), dim3(256), 0, st, pm, ps, nb, I, c->lse, blk, c->scal);
    hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(1024), 0, st, blk, nfin, c->scal + SC_TOTAL);
    const int ncb = (I + kApplyCols - 1) / kApplyCols;
    const dim3 ga(ncb, (S + kApplyRows - 1) / kApplyRows);
    double *nzpart = Tpart + (int64_t)ncb * S;  // one non-zero count per apply workgroup
    if (c->mfma)
        hipLaunchKernelGGL(k_sm_apply<1>, ga, dim3(256), 0, st, c->R, (int64_t)c->Ip, S, I, c->lse, c->Ph, c->Pl,
                           Tpart, nzpart, c->opt.scaling ? c->iphi : nullptr, c->pdirty, p_segments(c),
                           fused ? pm : nullptr, fused && c->logits_partial ? c->blk : nullptr, c->chunk,
                           c->chunkp / 256);
    else
        hipLaunchKernelGGL(k_sm_apply<0>, ga, dim3(256), 0, st, c->R, (int64_t)c->Ip, S, I, c->lse, nullptr, nullptr,
                           Tpart, nullptr, nullptr, nullptr, 0, nullptr, nullptr, 1, 1);
    hipLaunchKernelGGL(k_T_final, dim3((c->Sp + 255) / 256), dim3(256), 0, st, Tpart, ncb, S, c->Sp, c->T);
    if (c->mfma) hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(1024), 0, st, nzpart, (int)(ga.x * ga.y), c->scal + SC_NNZ);
    return hipGetLastError();
}

}  // namespace emc
