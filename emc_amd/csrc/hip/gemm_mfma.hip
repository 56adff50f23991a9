// MFMA contractions for CDNA4 (gfx950): C[m, n] = sum_k sum_{(a,b)} A_a[m, k] B_b[n, k]
// with A and B given as BF16 terms whose sum carries the operand (A_0 = hi, A_1 = next 8 bits, ...)
// and FP32 accumulation in the matrix cores.
//
//   logL   : A = centred log-slices (NA = 2 or 3 terms) [rows = rotations], B = counts
//            [rows = patterns] (NB = 1, or 2 when a count exceeds 256), k = active pixels.
//            Epilogue: R[s, i] = acc + bias_s.
//   M-step : A = responsibilities hi+lo [rows = rotations], B = counts^T [rows = pixels],
//            k = patterns.  Epilogue: store U, or scatter U straight into the model (fused).
//
// Term pairs (a, b) with a + b < NA are accumulated (the dropped pairs are below 2^-8NA of the
// product).  Tiling: 128 x 128 macro tile, 256 threads = 4 wavefronts of 64 lanes, each wave a
// 64 x 64 block of 2 x 2 v_mfma_f32_32x32x16_bf16 tiles; BK = 64; operands staged in LDS with
// rows padded to 72 BF16 (144 B) so the 16-byte fragment reads of 16 lanes hit 64 distinct
// banks; the next k-tile is fetched into registers while the current one is multiplied.
// The fused-scatter epilogue stages the accumulators through LDS, 64 rows at a time.
#include <algorithm>
#include <type_traits>

#include "../emc_internal.h"
#include "geom.h"

namespace emc {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

constexpr int BM = 128, BN = 128, BK = 64, LDK = BK + 8;

enum { EPI_LOGL = 0, EPI_STORE = 1, EPI_SCATTER = 2 };

struct EpiArgs {
    float *C;            // logits / U output
    int64_t ldc;
    int m_valid, n_valid;
    const float *bias;   // [rows] (logL)
    // fused scatter
    const double *T;     // [rows] slice sums
    const float *rotm;   // [rows][12]
    const float *axyz;   // [3][Pp]
    int Pp, Pa, L, interp;
    float *Mnum, *Wt;
    // softmax partials of the logL epilogue (large-tile kernel): per 256-row band and pattern,
    // the band maximum pm and sum of exp(R - pm) ps; band = band0 + m-tile
    float *pm;
    double *ps;
    int64_t band0, ldp;
    // the DMA logL kernel's store skipping (null: every logit is stored): per pattern the running
    // maximum of the band maxima (ord_of form), and the count of columns left out
    int32_t *runmax;
    double *skipped;
    int16_t *blk;  // [band][ldp] where each column's logits went (null: every logit in place)
    // device-side choice between this kernel and the sparse M-step: the whole grid returns at
    // once when *gate <= gate_thr (the responsibilities' non-zero count is small)
    const double *gate;
    double gate_thr;
    bool accumulate;  // store epilogue: C += acc instead of C = acc
    int tile_rows;    // large-tile kernel: row panels per tile group (0 = kTileRows)
};

__device__ __forceinline__ void scatter_one(const EpiArgs &e, const float *m, float Ts, int j, float u)
{
    if (!(u >= 0.0f) || j >= e.Pa) return;
    float ux, uy, uz;
    rot_point(m, e.axyz[j], e.axyz[e.Pp + j], e.axyz[2 * e.Pp + j], model_c(e.L), ux, uy, uz);
    const int L = e.L;
    if (e.interp == EMC_INTERP_NEAREST) {
        const int vx = nearest(ux), vy = nearest(uy), vz = nearest(uz);
        if (vx >= 0 && vx < L && vy >= 0 && vy < L && vz >= 0 && vz < L) {
            const size_t v = ((size_t)vz * L + vy) * L + vx;
            atomicAdd(e.Mnum + v, u);
            atomicAdd(e.Wt + v, Ts);
        }
    } else {
        int bx, by, bz;
        float fx, fy, fz;
        if (!tri_axis(ux, L, bx, fx) || !tri_axis(uy, L, by, fy) || !tri_axis(uz, L, bz, fz)) return;
        const float g[2][3] = {{1.0f - fx, 1.0f - fy, 1.0f - fz}, {fx, fy, fz}};
        const size_t base = ((size_t)bz * L + by) * L + bx;
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const float w = g[dx][0] * g[dy][1] * g[dz][2];
                    const size_t v = base + ((size_t)dz * L + dy) * L + dx;
                    atomicAdd(e.Mnum + v, w * u);
                    atomicAdd(e.Wt + v, w * Ts);
                }
    }
}

template <int NA, int NB, int EPI>
__global__ __launch_bounds__(256) void k_gemm_mfma(const bf16_t *__restrict__ A0, const bf16_t *__restrict__ A1,
                                                   const bf16_t *__restrict__ A2, int64_t lda,
                                                   const bf16_t *__restrict__ B0, const bf16_t *__restrict__ B1,
                                                   int64_t ldb, int Kd, EpiArgs e)
{
    if (e.gate && *e.gate <= e.gate_thr) return;
    constexpr int kTileBytes = (NA + NB) * BM * LDK * 2;
    static_assert(kTileBytes >= 64 * (BN + 4) * 4, "LDS too small for the epilogue staging");
    __shared__ __attribute__((aligned(16))) unsigned char smem[kTileBytes];
    bf16_t(*As)[BM][LDK] = reinterpret_cast<bf16_t(*)[BM][LDK]>(smem);
    bf16_t(*Bs)[BN][LDK] = reinterpret_cast<bf16_t(*)[BN][LDK]>(smem + NA * BM * LDK * 2);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;

    // 128 rows x 64 k x 2 B = 1024 chunks of 16 B per operand term; 4 per thread
    const int lrow = tid >> 3, lcol = (tid & 7) * 8;
    const bf16_t *ga[NA], *gb[NB];
    ga[0] = A0 + (size_t)(m0 + lrow) * lda + lcol;
    if (NA > 1) ga[1] = A1 + (size_t)(m0 + lrow) * lda + lcol;
    if (NA > 2) ga[NA - 1] = A2 + (size_t)(m0 + lrow) * lda + lcol;
    gb[0] = B0 + (size_t)(n0 + lrow) * ldb + lcol;
    if (NB > 1) gb[NB - 1] = B1 + (size_t)(n0 + lrow) * ldb + lcol;

    u32x4 ra[NA][4], rb[NB][4];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int h = 0; h < 4; ++h) ra[a][h] = *reinterpret_cast<const u32x4 *>(ga[a] + (size_t)32 * h * lda);
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int h = 0; h < 4; ++h) rb[b][h] = *reinterpret_cast<const u32x4 *>(gb[b] + (size_t)32 * h * ldb);

    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int nk = Kd / BK;
    const int fr = lane & 31, fk = (lane >> 5) * 8;
    for (int kt = 0; kt < nk; ++kt) {
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int h = 0; h < 4; ++h) *reinterpret_cast<u32x4 *>(&As[a][lrow + 32 * h][lcol]) = ra[a][h];
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int h = 0; h < 4; ++h) *reinterpret_cast<u32x4 *>(&Bs[b][lrow + 32 * h][lcol]) = rb[b][h];
        __syncthreads();
        if (kt + 1 < nk) {
            const int k1 = (kt + 1) * BK;
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int h = 0; h < 4; ++h)
                    ra[a][h] = *reinterpret_cast<const u32x4 *>(ga[a] + (size_t)32 * h * lda + k1);
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int h = 0; h < 4; ++h)
                    rb[b][h] = *reinterpret_cast<const u32x4 *>(gb[b] + (size_t)32 * h * ldb + k1);
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            bf16x8 fa[NA][2], fb[NB][2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
#pragma unroll
                for (int a = 0; a < NA; ++a)
                    fa[a][t] = *reinterpret_cast<const bf16x8 *>(&As[a][wm + t * 32 + fr][kk + fk]);
#pragma unroll
                for (int b = 0; b < NB; ++b)
                    fb[b][t] = *reinterpret_cast<const bf16x8 *>(&Bs[b][wn + t * 32 + fr][kk + fk]);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int a = NA - 1; a >= 0; --a)
#pragma unroll
                        for (int b = NB - 1; b >= 0; --b)
                            if (a + b < NA)
                                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[a][i], fb[b][j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }

    // accumulator layout: lane holds column n = lane%32 of each 32x32 tile and rows
    // 8*(r/4) + 4*(lane/32) + r%4, r = 0..15
    if (EPI != EPI_SCATTER) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (m >= e.m_valid) continue;
                const float add = EPI == EPI_LOGL ? e.bias[m] : 0.0f;
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int n = n0 + wn + j * 32 + (lane & 31);
                    if (n < e.n_valid) e.C[(size_t)m * e.ldc + n] = acc[i][j][r] + add;
                }
            }
        return;
    }

    // fused scatter: stage one 64-row half of the tile in LDS, then every thread scatters a
    // strided share of it (one compact loop instead of 64 unrolled scatter bodies)
    float(*Cs)[BN + 4] = reinterpret_cast<float(*)[BN + 4]>(smem);
    for (int half = 0; half < 2; ++half) {
        if ((wave >> 1) == half) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int r = 0; r < 16; ++r)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        Cs[i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)][wn + j * 32 + (lane & 31)] = acc[i][j][r];
        }
        __syncthreads();
        for (int row = wave; row < 64; row += 4) {
            const int m = m0 + half * 64 + row;
            if (m >= e.m_valid) continue;
            const double Td = e.T[m];
            if (!(Td > 1e-10)) continue;
            float rm[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) rm[k] = e.rotm[12 * (size_t)m + k];
            for (int col = lane; col < BN; col += 64) scatter_one(e, rm, (float)Td, n0 + col, Cs[row][col]);
        }
        __syncthreads();
    }
}

template <int NA, int NB, int EPI>
hipError_t run(dim3 grid, hipStream_t st, const bf16_t *A0, const bf16_t *A1, const bf16_t *A2, int64_t lda,
               const bf16_t *B0, const bf16_t *B1, int64_t ldb, int Kd, const EpiArgs &e)
{
    hipLaunchKernelGGL((k_gemm_mfma<NA, NB, EPI>), grid, dim3(256), 0, st, A0, A1, A2, lda, B0, B1, ldb, Kd, e);
    return hipGetLastError();
}

// ---- large-tile variant ---------------------------------------------------------------------
// 256 x BN2 macro tile, 512 threads = 8 wavefronts (4 along M x 2 along N), each wave a 64 x BN2/2
// block of 2 x (BN2/64) 32x32 tiles, so every A fragment (NA terms) feeds BN2/64 MFMAs and every
// B fragment NA * 2.  BK = 32; two LDS stages (rows padded to 40 BF16 = 80 B: the 16-byte
// fragment reads of 16 lanes cover all 64 banks once), one barrier per k-step; the global loads
// run two k-steps ahead through two register sets (one step ahead left the load latency exposed:
// 6.15 -> 5.6 ms at the 64^3 / 25,680 / 20,000 logL shape).
// Workgroups are renumbered so that each XCD (round-robin dispatch: workgroup b runs on XCD b % 8)
// works on a contiguous band of tiles, which share A row panels in that XCD's L2.
constexpr int BM2 = 256, BK2 = 32, LDK2 = BK2 + 8;
constexpr int kTileRows = 1;  // default tile grouping of the large-tile kernel

// Counts stored as bytes (exact to 255) widen to BF16 on the way to LDS: 8 bytes -> 8 BF16.
__device__ __forceinline__ u32x4 widen_u8x8(u32x2 v)
{
    u32x4 o;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t w = v[k >> 1] >> ((k & 1) * 16);
        const uint32_t f0 = __builtin_bit_cast(uint32_t, (float)(w & 0xffu));
        const uint32_t f1 = __builtin_bit_cast(uint32_t, (float)((w >> 8) & 0xffu));
        o[k] = (f0 >> 16) | (f1 & 0xffff0000u);  // integers < 256 have zero low mantissa bits
    }
    return o;
}

// Order-preserving integer form of a float (signed compare = float compare, NaN aside), for the
// running per-pattern maxima kept with integer atomicMax; kOrdNegInf is the form of -inf.
__device__ __forceinline__ int32_t ord_of(float x)
{
    const int32_t b = __float_as_int(x);
    return b >= 0 ? b : b ^ 0x7fffffff;
}
__device__ __forceinline__ float float_of_ord(int32_t k) { return __int_as_float(k >= 0 ? k : k ^ 0x7fffffff); }
constexpr int32_t kOrdNegInf = (int32_t)0x807fffff;

// A band's logits of a pattern are not stored when the band's largest one lies more than this many
// nats below a largest logit already found for the pattern: every responsibility in the band is then
// below exp(-128) of the pattern's largest one, and the softmax, which treats every band more than
// kSkipApplyNats below the pattern's lse as -inf, computes P = 0 there from the band maxima alone.
constexpr double kSkipStoreNats = 128.0;

// logL epilogue of the 256-row large-tile kernels (8 waves: 4 along M x 2 along N, each a 64-row
// by BN2/2-column block of 2 x TN 32x32 tiles).  Accumulator layout: lane holds column lane % 32
// of each 32x32 tile and rows 8*(r/4) + 4*(lane/32) + r%4, r = 0..15.  R = acc + bias_s.  The
// softmax partials of this 256-row band (e.pm and e.ps: the large-tile logL kernels run only with
// the softmax fused) restate k_sm_partial's rules: column maxima from the -1e10 floor, over the lane
// pair, then over the 4 row-waves through LDS; then the sums of exp(R - max), where a term with
// R - max <= -1e10 (or not a number) adds nothing.  Each lane sums its 32 terms of a column in FP32
// on the hardware exp2 (the exponent floored at -1e10, whose exp2 is 0); the lane and wave partials
// are added in FP64.
// The logits are stored after the maxima are known.  Without e.blk every logit is stored in place
// (row-major R).  With e.blk (the iteration's store skipping): a column whose band maximum lies
// kSkipStoreNats below rmk[column of the tile] (running maxima read when the workgroup started:
// lower bounds of the pattern's largest logit) is not stored; in a tile inside the valid rows and
// columns the kept columns are packed, in column order, as blocks of 256 consecutive floats (the
// column's rows) in the tile's own rows of R (block k in row m0 + k, columns n0 .. n0 + 255), so the
// HBM writes are whole lines however few columns are kept; the other tiles store their kept columns
// in place.  e.blk[band][pattern] records the block (>= 0), -2 (in place) or -1 (not stored); every
// band maximum is merged into e.runmax, and e.skipped counts the columns left out.
// A tile inside the valid rows and columns (every tile but the last row band and column tile) takes
// a path without bounds checks.  lds: 48 BN2 + 1056 bytes, free (the main loop's stages).  bias_lds:
// the tile's 256 row biases, copied to LDS by the kernel (0 past m_valid), so no vector-memory wait
// holds up the logit stores (on gfx9 a wait for a load also waits for every store issued before it)
// and no bias stays in registers between the passes.
template <int TN, int BN2, bool kFull>
__device__ __forceinline__ void logl_epi(const f32x16 (&acc)[2][TN], const EpiArgs &e, int m0, int n0, int wm, int wn,
                                         void *lds, const float *bias_lds, const int32_t *rmk)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    wm = __builtin_amdgcn_readfirstlane(wm);
    wn = __builtin_amdgcn_readfirstlane(wn);
    const int hr = 4 * (lane >> 5), cl = lane & 31;
    const uint32_t lo = (uint32_t)(hr * e.ldc + cl);  // lane offset from a row's base
    // bias of accumulator row r of tile i; -inf past m_valid (adds to neither the maximum nor the sum)
    auto bias = [&](int i, int r) {
        const int mt = wm + i * 32 + 8 * (r >> 2) + (r & 3) + hr;
        return kFull || m0 + mt < e.m_valid ? bias_lds[mt] : -INFINITY;
    };
    float cmax[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) cmax[j] = -1.0e10f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const float add = bias(i, r);
#pragma unroll
            for (int j = 0; j < TN; ++j) cmax[j] = fmaxf(cmax[j], acc[i][j][r] + add);
        }
    float *red = reinterpret_cast<float *>(lds);            // [4][BN2] maxima
    double *reds = reinterpret_cast<double *>(lds) + 2 * BN2;  // [4][BN2] sums (after the maxima)
    int *cnt = reinterpret_cast<int *>(reinterpret_cast<unsigned char *>(lds) + 48 * BN2 + 1024);  // [2 TN]
    const int wrow = wave >> 1, half = wn / (BN2 / 2);
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        cmax[j] = fmaxf(cmax[j], __shfl_xor(cmax[j], 32, 64));
        if (lane < 32) red[wrow * BN2 + wn + j * 32 + lane] = cmax[j];
    }
    __syncthreads();
    bool keep[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const int c = wn + j * 32 + cl;
        cmax[j] = fmaxf(fmaxf(red[c], red[BN2 + c]), fmaxf(red[2 * BN2 + c], red[3 * BN2 + c]));
        keep[j] = !(rmk && (double)cmax[j] < (double)float_of_ord(rmk[c]) - kSkipStoreNats);
        if (!kFull) keep[j] = keep[j] && n0 + c < e.n_valid;
    }
    const bool compact = kFull && e.blk && BN2 == 256;  // a block is 256 floats of the tile's row
    // kept columns of each group of 32 (the same in every row-wave: keep depends on the column only)
    uint32_t km[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) km[j] = (uint32_t)__ballot(keep[j]);
    if (compact && wrow == 0 && lane == 0)
#pragma unroll
        for (int j = 0; j < TN; ++j) cnt[half * TN + j] = __popc(km[j]);
    if (rmk && wrow == 0) {  // the columns left out, counted once (row-wave 0, lanes 0-31)
        int nskip = 0;
#pragma unroll
        for (int j = 0; j < TN; ++j) nskip += __popc(~km[j] & (uint32_t)__ballot(n0 + wn + j * 32 + cl < e.n_valid));
        if (lane == 0 && nskip > 0) atomicAdd(e.skipped, (double)nskip);
    }
    constexpr float kLog2e = 1.4426950408889634f;
    float fs[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) fs[j] = 0.0f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const float add = bias(i, r);
#pragma unroll
            for (int j = 0; j < TN; ++j)
                fs[j] += __builtin_amdgcn_exp2f(fmaxf((acc[i][j][r] + add) - cmax[j], -1.0e10f) * kLog2e);
        }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        double s = (double)fs[j];
        s += __shfl_xor(s, 32, 64);
        if (lane < 32) reds[wrow * BN2 + wn + j * 32 + lane] = s;
    }
    __syncthreads();  // also publishes the kept-column counts
    const int64_t orow = (e.band0 + m0 / BM2) * e.ldp;  // this band's row of pm, ps and blk
    if (compact) {
        int off = 0;  // blocks of the groups before (half, j) in column order
#pragma unroll
        for (int g = 0; g < 2 * TN; ++g) off += g < half * TN ? cnt[g] : 0;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int rank = off + __popc(km[j] & ((1u << cl) - 1u));
            off += __popc(km[j]);
            if (wrow == 0 && lane < 32) e.blk[orow + n0 + wn + j * 32 + cl] = (int16_t)(keep[j] ? rank : -1);
            if (!keep[j]) continue;
            float *blk = e.C + (size_t)(m0 + rank) * e.ldc + n0;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int g = 0; g < 4; ++g) {  // rows 8 g + hr + 0..3 of tile i: 16 contiguous bytes
                    const int mr = wm + i * 32 + 8 * g + hr;
                    const float4 v = make_float4(acc[i][j][4 * g] + bias(i, 4 * g), acc[i][j][4 * g + 1] + bias(i, 4 * g + 1),
                                                 acc[i][j][4 * g + 2] + bias(i, 4 * g + 2),
                                                 acc[i][j][4 * g + 3] + bias(i, 4 * g + 3));
                    *reinterpret_cast<float4 *>(blk + mr) = v;
                }
        }
    } else {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            if (e.blk && wrow == 0 && lane < 32 && n0 + wn + j * 32 + cl < e.n_valid)
                e.blk[orow + n0 + wn + j * 32 + cl] = (int16_t)(keep[j] ? -2 : -1);
            if (!keep[j]) continue;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int mt = wm + i * 32 + 8 * (r >> 2) + (r & 3);  // tile row of lanes 0-31
                    if (!kFull && m0 + mt + hr >= e.m_valid) continue;
                    float *row = e.C + (size_t)(m0 + mt) * e.ldc + (n0 + wn);
                    row[lo + j * 32] = acc[i][j][r] + bias(i, r);
                }
        }
    }
    for (int c = tid; c < BN2; c += 512) {
        const int n = n0 + c;
        if (n >= e.n_valid) continue;
        const float mx = fmaxf(fmaxf(red[c], red[BN2 + c]), fmaxf(red[2 * BN2 + c], red[3 * BN2 + c]));
        const double sm = reds[c] + reds[BN2 + c] + reds[2 * BN2 + c] + reds[3 * BN2 + c];
        e.pm[orow + n] = mx;
        e.ps[orow + n] = sm;
        if (e.runmax && mx == mx) atomicMax(e.runmax + n, ord_of(mx));
    }
}

template <int TN, int BN2>
__device__ __forceinline__ void logl_epilogue(const f32x16 (&acc)[2][TN], const EpiArgs &e, int m0, int n0, int wm, int wn,
                                              void *lds, const float *bias_lds, const int32_t *rmk = nullptr)
{
    if (m0 + BM2 <= e.m_valid && n0 + BN2 <= e.n_valid) logl_epi<TN, BN2, true>(acc, e, m0, n0, wm, wn, lds, bias_lds, rmk);
    else logl_epi<TN, BN2, false>(acc, e, m0, n0, wm, wn, lds, bias_lds, rmk);
}

template <int NA, int EPI, int BN2, bool BU8>
__global__ __launch_bounds__(512) void k_gemm_mfma2(const bf16_t *__restrict__ A0, const bf16_t *__restrict__ A1,
                                                    const bf16_t *__restrict__ A2, int64_t lda,
                                                    const void *__restrict__ Bv, int64_t ldb, int Kd, int ntn,
                                                    int n_base, int gm, EpiArgs e)
{
    if (e.gate && *e.gate <= e.gate_thr) return;
    constexpr int TN = BN2 / 64;                       // 32-wide tiles per wave along N
    constexpr int BCH = BN2 * (BK2 / 8) / 512;         // 16-byte B chunks per thread per stage
    constexpr int STAGE = (NA * BM2 + BN2) * LDK2;     // BF16 elements per stage
    __shared__ __attribute__((aligned(16))) bf16_t smem[2 * STAGE];

    // XCD-aware renumbering (bijective for any grid size), then tiles in groups of gm row panels:
    // consecutive tiles t walk the gm rows of a group fastest, so the workgroups an XCD runs at
    // once share gm A panels and (their count / gm) B panels in its L2
    const int nwg = gridDim.x, b = blockIdx.x;
    const int base = nwg / 8, rem = nwg % 8, xcd = b % 8, loc = b / 8;
    const int t = xcd * base + min(xcd, rem) + loc;
    const int ntm = nwg / ntn, gi = t / (gm * ntn), gr = t % (gm * ntn);
    const int rows = min(gm, ntm - gi * gm);
    const int m0 = (gi * gm + gr % rows) * BM2, n0 = n_base + (gr / rows) * BN2;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * (BN2 / 2);
    // One 16-byte chunk per thread and 128-row half: k offset lcol of row lrow.  Lanes 4q..4q+3
    // hold the four chunks of one row, and the two rows of a lane octet are 4 apart (rows 8k + j
    // and 8k + j + 4): with 80-byte LDS rows the octet's stores then fall on 8 distinct 16-byte
    // bank groups (adjacent rows would share one).
    const int lrow = ((tid >> 5) << 3) + ((tid >> 3) & 3) + (((tid >> 2) & 1) << 2), lcol = (tid & 3) * 8;

    const bf16_t *ga[NA];
    ga[0] = A0 + (size_t)(m0 + lrow) * lda + lcol;
    if (NA > 1) ga[1] = A1 + (size_t)(m0 + lrow) * lda + lcol;
    if (NA > 2) ga[NA - 1] = A2 + (size_t)(m0 + lrow) * lda + lcol;
    // B element size 1 (bytes) or 2 (BF16); one chunk = 8 consecutive k
    const unsigned char *gb = reinterpret_cast<const unsigned char *>(Bv) + ((size_t)(n0 + lrow) * ldb + lcol) * (BU8 ? 1 : 2);

    // two register sets: the global loads run two k-steps ahead of the multiplication
    u32x4 ra[2][NA][2], rb[2][BCH];
    u32x2 rb8[2][BCH];
    auto gload = [&](int set, int k0) {
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int h = 0; h < 2; ++h) ra[set][a][h] = *reinterpret_cast<const u32x4 *>(ga[a] + (size_t)128 * h * lda + k0);
#pragma unroll
        for (int h = 0; h < BCH; ++h) {
            if (BU8) rb8[set][h] = *reinterpret_cast<const u32x2 *>(gb + (size_t)128 * h * ldb + k0);
            else rb[set][h] = *reinterpret_cast<const u32x4 *>(gb + ((size_t)128 * h * ldb + k0) * 2);
        }
    };
    auto sstore = [&](int set, int stage) {
        bf16_t *S = smem + stage * STAGE;
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int h = 0; h < 2; ++h)
                *reinterpret_cast<u32x4 *>(S + ((size_t)a * BM2 + lrow + 128 * h) * LDK2 + lcol) = ra[set][a][h];
#pragma unroll
        for (int h = 0; h < BCH; ++h)
            *reinterpret_cast<u32x4 *>(S + ((size_t)NA * BM2 + lrow + 128 * h) * LDK2 + lcol) =
                BU8 ? widen_u8x8(rb8[set][h]) : rb[set][h];
    };

    f32x16 acc[2][TN];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int fr = lane & 31, fk = (lane >> 5) * 8;
    auto compute = [&](int stage) {
        const bf16_t *S = smem + stage * STAGE;
#pragma unroll
        for (int kk = 0; kk < BK2; kk += 16) {
            bf16x8 fa[NA][2], fb[TN];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int a = 0; a < NA; ++a)
                    fa[a][i] = *reinterpret_cast<const bf16x8 *>(S + ((size_t)a * BM2 + wm + i * 32 + fr) * LDK2 + kk + fk);
#pragma unroll
            for (int j = 0; j < TN; ++j)
                fb[j] = *reinterpret_cast<const bf16x8 *>(S + ((size_t)NA * BM2 + wn + j * 32 + fr) * LDK2 + kk + fk);
#pragma unroll
            for (int a = NA - 1; a >= 0; --a)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[a][i], fb[j], acc[i][j], 0, 0, 0);
        }
    };

    // Step k is multiplied from LDS stage k % 2 while the loads of step k + 2 are in flight into
    // register set k % 2 (stored at the end of step k - 1) and those of step k + 1 land in set
    // (k + 1) % 2, which is written to the other stage at the end of step k.  The loop is unrolled
    // by two so every set and stage index is a constant.
    // The loads are unconditional (past the last step they re-read it) so the compiler can count
    // the loads in flight and wait for one register set only, and a scheduling barrier keeps them
    // at the top of the step (the scheduler otherwise sinks them below the multiplications).
    const int nk = Kd / BK2;
    gload(0, 0);
    gload(1, min(1, nk - 1) * BK2);
    sstore(0, 0);
    __syncthreads();
    int kt = 0;
    for (; kt + 1 < nk; kt += 2) {
        gload(0, min(kt + 2, nk - 1) * BK2);
        __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the multiplications
        compute(0);
        sstore(1, 1);
        __syncthreads();
        gload(1, min(kt + 3, nk - 1) * BK2);
        __builtin_amdgcn_sched_barrier(0);
        compute(1);
        sstore(0, 0);  // past the last step: a copy of it that nothing reads
        __syncthreads();
    }
    if (kt < nk) {  // odd step count: the last step was stored in stage 0
        compute(0);
        __syncthreads();  // the epilogue reuses the LDS
    }

    if constexpr (EPI == EPI_LOGL) {
        float *bias_l = reinterpret_cast<float *>(reinterpret_cast<unsigned char *>(smem) + 48 * BN2);
        if (tid < BM2) bias_l[tid] = m0 + tid < e.m_valid ? e.bias[m0 + tid] : 0.0f;
        __syncthreads();
        logl_epilogue<TN, BN2>(acc, e, m0, n0, wm, wn, smem, bias_l);
    } else {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (m >= e.m_valid) continue;
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    const int n = n0 + wn + j * 32 + (lane & 31);
                    if (n < e.n_valid) {
                        float *cp = e.C + (size_t)m * e.ldc + n;
                        *cp = e.accumulate ? *cp + acc[i][j][r] : acc[i][j][r];
                    }
                }
            }
    }
}

// ---- contractions with DMA operand staging ---------------------------------------------------
// Two-term A (log W or P, hi + lo) x one-term BF16 B (counts), 256 x 256 macro tile, 8 wavefronts
// as in k_gemm_mfma2, BK2 = 32; the logL epilogue (k_logl_dma) or the store epilogue (k_store_dma:
// dense M-step, chi^2 sums).  The operands go from global memory straight into LDS through the
// vector-memory DMA path (global_load_lds_dwordx4: lane l of a wave instruction lands at
// M0 + 16 l), so no register staging and no LDS store instructions.  LDS rows are 64 B (32 BF16)
// without padding; the four 16-byte chunks of row r sit at chunk ^ lds_swz(r), which spreads the
// fragment reads of 8 (or 16) consecutive rows over distinct bank groups.  Three stages: the loads
// of step k + 2 are issued during step k.  Each step reads the fragments of its second k-half
// while the first half multiplies, and those of the next step's first half (after the barrier)
// while the second half multiplies.
// Measured for the logL at 25,856 x 20,224 x 3,200 against the register-staged k_gemm_mfma2: 5.70
// vs 5.90 ms and 5.86 vs 6.28 ms on two boxes, bit-identical logits and band partials.
constexpr int kDmaStages = 3;

__device__ __forceinline__ int lds_swz(int r) { return ((r >> 2) & 1) | ((((r >> 1) ^ (r >> 3)) & 1) << 1); }

// One DMA load of 16 bytes per lane into LDS at the wave-uniform byte address lds.  Written as
// inline asm: with the builtin, the compiler's LDS-DMA tracking puts a vmcnt(0) in front of every
// LDS read, which would wait for the loads of later stages too; the kernel waits with explicit
// vmcnt counts instead.  The instruction has no register result; the s_nop is the one wait state
// required between writing M0 and an LDS DMA.  M0 is a reserved register to the compiler; the DMA
// kernels' code uses it nowhere else (clang warns about the clobber, which is silenced here).
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void lds_dma16(const void *g, uint32_t lds)
{
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(g), "s"(lds) : "memory", "m0");
}
#pragma clang diagnostic pop

template <int EPI>
__device__ __forceinline__ void mfma_dma_body(const bf16_t *__restrict__ A0, const bf16_t *__restrict__ A1,
                                              int64_t lda, const bf16_t *__restrict__ B16, int64_t ldb, int Kd,
                                              int ntn, int gm, const EpiArgs &e)
{
    static_assert(EPI == EPI_LOGL || EPI == EPI_STORE, "logL or store epilogue");
    if (e.gate && *e.gate <= e.gate_thr) return;  // the sparse M-step runs instead
    constexpr int NA = 2, BN2 = 256, TN = BN2 / 64;
    constexpr int OPB = 256 * BK2 * 2;        // bytes of one operand per stage
    constexpr int STAGEB = (NA + 1) * OPB;    // 48 KB
    constexpr int LPW = (NA + 1) * 16 / 8;    // DMA instructions per wave per step (16 rows each)
    constexpr uint32_t kVm0 = 0x0F70;         // s_waitcnt encoding: vmcnt(0), lgkmcnt and expcnt not waited
    __shared__ __attribute__((aligned(1024))) unsigned char smem[kDmaStages * STAGEB];
    __shared__ float bias_s[BM2];
    __shared__ int32_t rm_s[BN2];

    // Tile order as in k_gemm_mfma2 (XCD x = b % 8 takes its contiguous share of tiles in the order
    // of b / 8), except that a last row band holding padded rows is left to the end of every XCD's
    // share (its ntn tiles split over the XCDs): the waves of its padded rows skip their products,
    // so those tiles cost about half, and each XCD's final, partly filled round is made of them.
    const int nwg = gridDim.x, b = blockIdx.x;
    const int base = nwg / 8, rem = nwg % 8, xcd = b % 8, loc = b / 8;
    const int ntm = nwg / ntn;
    int m0, n0;
    if (e.m_valid < ntm * BM2 && ntm > 1 && ntn >= 8) {  // then every XCD has room for its share
        const int pc = ntn / 8 + (xcd < ntn % 8 ? 1 : 0), pstart = xcd * (ntn / 8) + min(xcd, ntn % 8);
        const int fc = base + (xcd < rem ? 1 : 0) - pc, fstart = xcd * base + min(xcd, rem) - pstart;
        if (loc < fc) {
            const int t = fstart + loc, gi = t / (gm * ntn), gr = t % (gm * ntn);
            const int rows = min(gm, ntm - 1 - gi * gm);
            m0 = (gi * gm + gr % rows) * BM2;
            n0 = (gr / rows) * BN2;
        } else {
            m0 = (ntm - 1) * BM2;
            n0 = (pstart + loc - fc) * BN2;
        }
    } else {
        const int t = xcd * base + min(xcd, rem) + loc;
        const int gi = t / (gm * ntn), gr = t % (gm * ntn);
        const int rows = min(gm, ntm - gi * gm);
        m0 = (gi * gm + gr % rows) * BM2;
        n0 = (gr / rows) * BN2;
    }

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * (BN2 / 2);
    const bool live = m0 + wm < e.m_valid;  // a wave whose 64 rows are all padding multiplies nothing
    // DMA instruction q of this wave: operand op (A hi, A lo, B), rows 16 rb .. 16 rb + 15; lane l
    // fetches logical chunk (l & 3) ^ lds_swz(r) of row r = 16 rb + l / 4
    const bf16_t *src[LPW];
    int dst[LPW];
#pragma unroll
    for (int q = 0; q < LPW; ++q) {
        const int g = wave * LPW + q, op = g >> 4, rb = g & 15;
        const int r = rb * 16 + (lane >> 2), c = (lane & 3) ^ lds_swz(r);
        const bf16_t *row = op == 0 ? A0 + (size_t)(m0 + r) * lda
                          : op == 1 ? A1 + (size_t)(m0 + r) * lda
                                    : B16 + (size_t)(n0 + r) * ldb;
        src[q] = row + c * 8;
        dst[q] = op * OPB + rb * 1024;
    }
    const uint32_t lds0 = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) unsigned char *)smem;
    auto issue = [&](int step, int stage) {
#pragma unroll
        for (int q = 0; q < LPW; ++q) lds_dma16(src[q] + step * BK2, lds0 + stage * STAGEB + dst[q]);
    };

    f32x16 acc[2][TN];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
    const int fr = lane & 31, sw = lds_swz(fr & 15), c0 = lane >> 5;
    struct Frag {
        bf16x8 a[NA][2], b[TN];
    };
    auto rd = [&](Frag &F, int stage, int kk) {  // fragments of k-half kk (0 or 16) of a stage
        const unsigned char *S = smem + stage * STAGEB;
        const int co = ((c0 + kk / 8) ^ sw) * 16;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int a = 0; a < NA; ++a)
                F.a[a][i] = *reinterpret_cast<const bf16x8 *>(S + a * OPB + (wm + i * 32 + fr) * 64 + co);
#pragma unroll
        for (int j = 0; j < TN; ++j) F.b[j] = *reinterpret_cast<const bf16x8 *>(S + NA * OPB + (wn + j * 32 + fr) * 64 + co);
    };
    auto mm = [&](const Frag &F) {  // lo term first, as in k_gemm_mfma2
#pragma unroll
        for (int a = NA - 1; a >= 0; --a)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(F.a[a][i], F.b[j], acc[i][j], 0, 0, 0);
    };

    // stage of step k: k % 3.  The loads of step k + 2 go to the stage of step k - 1, whose reads
    // all completed before the barrier in step k - 1.  The barrier in step k follows the wait for
    // this wave's loads of step k + 1 (issued in step k - 1): LPW newer instructions may remain.
    const int nk = Kd / BK2;
    Frag F0, F1;
    const float bv = EPI == EPI_LOGL && tid < BM2 && m0 + tid < e.m_valid ? e.bias[m0 + tid] : 0.0f;
    // running maxima of the tile's columns for the store skipping, read now (kept in LDS): a value
    // read earlier is still a lower bound of the pattern's largest logit when the epilogue uses it
    const int32_t rv = EPI == EPI_LOGL && e.runmax && tid < BN2 && n0 + tid < e.n_valid ? e.runmax[n0 + tid] : kOrdNegInf;
    issue(0, 0);
    if (nk > 1) issue(1, 1);
    if (nk > 1) __builtin_amdgcn_s_waitcnt(kVm0 | LPW);
    else __builtin_amdgcn_s_waitcnt(kVm0);
    if (tid < BM2) bias_s[tid] = bv;
    if (tid < BN2) rm_s[tid] = rv;
    __syncthreads();
    // the k loop; a wave of padded rows (LIVE false) only issues its share of the loads and keeps
    // the barriers
    auto kloop = [&](auto live_tag) {
        constexpr bool LIVE = decltype(live_tag)::value;
        if (LIVE) rd(F0, 0, 0);
        int k = 0;
        for (; k + 2 < nk; ++k) {
            if (LIVE) rd(F1, k % kDmaStages, 16);
            issue(k + 2, (k + 2) % kDmaStages);
            __builtin_amdgcn_sched_barrier(0);
            if (LIVE) mm(F0);
            __builtin_amdgcn_s_waitcnt(kVm0 | LPW);
            __syncthreads();
            if (LIVE) rd(F0, (k + 1) % kDmaStages, 0);
            __builtin_amdgcn_sched_barrier(0);
            if (LIVE) mm(F1);
        }
        if (k + 1 < nk) {
            if (LIVE) rd(F1, k % kDmaStages, 16);
            __builtin_amdgcn_sched_barrier(0);
            if (LIVE) mm(F0);
            __builtin_amdgcn_s_waitcnt(kVm0);
            __syncthreads();
            if (LIVE) rd(F0, (k + 1) % kDmaStages, 0);
            __builtin_amdgcn_sched_barrier(0);
            if (LIVE) mm(F1);
            ++k;
        }
        if (LIVE) {
            rd(F1, k % kDmaStages, 16);
            mm(F0);
            mm(F1);
        }
    };
    if (live) kloop(std::true_type{});
    else kloop(std::false_type{});
    __syncthreads();  // the epilogue reuses the LDS
    if constexpr (EPI == EPI_LOGL) {
        logl_epilogue<TN, BN2>(acc, e, m0, n0, wm, wn, smem, bias_s, e.runmax ? rm_s : nullptr);
    } else {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (m >= e.m_valid) continue;
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    const int n = n0 + wn + j * 32 + (lane & 31);
                    if (n < e.n_valid) {
                        float *cp = e.C + (size_t)m * e.ldc + n;
                        *cp = e.accumulate ? *cp + acc[i][j][r] : acc[i][j][r];
                    }
                }
            }
    }
}

// logL: R = log W . K^T + bias with the softmax band partials; store: C (=|+=) A . B^T (the dense
// M-step U = P^T . K and the chi^2 contractions).
__global__ __launch_bounds__(512) void k_logl_dma(const bf16_t *__restrict__ A0, const bf16_t *__restrict__ A1,
                                                  int64_t lda, const bf16_t *__restrict__ B16, int64_t ldb, int Kd,
                                                  int ntn, int gm, EpiArgs e)
{
    mfma_dma_body<EPI_LOGL>(A0, A1, lda, B16, ldb, Kd, ntn, gm, e);
}

__global__ __launch_bounds__(512) void k_store_dma(const bf16_t *__restrict__ A0, const bf16_t *__restrict__ A1,
                                                   int64_t lda, const bf16_t *__restrict__ B16, int64_t ldb, int Kd,
                                                   int ntn, int gm, EpiArgs e)
{
    mfma_dma_body<EPI_STORE>(A0, A1, lda, B16, ldb, Kd, ntn, gm, e);
}

template <int NA, int EPI, int BN2, bool BU8>
hipError_t run2(int ntm, int ntn, int n_base, hipStream_t st, const bf16_t *A0, const bf16_t *A1, const bf16_t *A2,
                int64_t lda, const void *B, int64_t ldb, int Kd, const EpiArgs &e)
{
    const int gm = std::max(1, std::min(e.tile_rows > 0 ? e.tile_rows : kTileRows, ntm));
    hipLaunchKernelGGL((k_gemm_mfma2<NA, EPI, BN2, BU8>), dim3(ntm * ntn), dim3(512), 0, st, A0, A1, A2, lda, B, ldb,
                       Kd, ntn, n_base, gm, e);
    return hipGetLastError();
}

// Large-tile kernel for a single-term B operand (BF16, or bytes when B8 is given).  One or two A
// terms: 256-wide column tiles, plus one 128-wide launch for the remainder when N % 256 = 128;
// three A terms: 128-wide tiles (LDS).
template <int EPI, bool U8>
hipError_t run_large_t(int M, int N, hipStream_t st, int na, const bf16_t *A0, const bf16_t *A1, const bf16_t *A2,
                       int64_t lda, const void *B, int64_t ldb, int Kd, const EpiArgs &e)
{
    const int ntm = M / BM2;
    if (na == 3) return run2<3, EPI, 128, U8>(ntm, N / 128, 0, st, A0, A1, A2, lda, B, ldb, Kd, e);
    const int n256 = N / 256;
    hipError_t err = hipSuccess;
    if (na == 1) {
        if (n256 > 0) err = run2<1, EPI, 256, U8>(ntm, n256, 0, st, A0, A1, A2, lda, B, ldb, Kd, e);
        if (err == hipSuccess && N % 256)
            err = run2<1, EPI, 128, U8>(ntm, 1, n256 * 256, st, A0, A1, A2, lda, B, ldb, Kd, e);
        return err;
    }
    if (n256 > 0) err = run2<2, EPI, 256, U8>(ntm, n256, 0, st, A0, A1, A2, lda, B, ldb, Kd, e);
    if (err == hipSuccess && N % 256) err = run2<2, EPI, 128, U8>(ntm, 1, n256 * 256, st, A0, A1, A2, lda, B, ldb, Kd, e);
    return err;
}

template <int EPI>
hipError_t run_large(int M, int N, hipStream_t st, int na, const bf16_t *A0, const bf16_t *A1, const bf16_t *A2,
                     int64_t lda, const bf16_t *B16, const uint8_t *B8, int64_t ldb, int Kd, const EpiArgs &e)
{
    if (B8) return run_large_t<EPI, true>(M, N, st, na, A0, A1, A2, lda, B8, ldb, Kd, e);
    return run_large_t<EPI, false>(M, N, st, na, A0, A1, A2, lda, B16, ldb, Kd, e);
}

}  // namespace

// Two-term A x one-term B with the DMA kernels over the 256-wide column tiles (BF16 B operand),
// plus the 128-wide register-staged remainder launch when N % 256 = 128.
template <int EPI>
hipError_t run_dma(int M, int N, hipStream_t st, const bf16_t *A0, const bf16_t *A1, int64_t lda,
                   const bf16_t *B16, const uint8_t *B8, int64_t ldb, int Kd, const EpiArgs &e)
{
    const int ntm = M / BM2, n256 = N / 256;
    hipError_t err = hipSuccess;
    if (n256 > 0) {
        const int gm = std::max(1, std::min(e.tile_rows > 0 ? e.tile_rows : kTileRows, ntm));
        hipLaunchKernelGGL(EPI == EPI_LOGL ? k_logl_dma : k_store_dma, dim3(ntm * n256), dim3(512), 0, st, A0, A1, lda,
                           B16, ldb, Kd, n256, gm, e);
        err = hipGetLastError();
    }
    if (err == hipSuccess && N % 256) {
        if (B8) err = run2<2, EPI, 128, true>(ntm, 1, n256 * 256, st, A0, A1, nullptr, lda, B8, ldb, Kd, e);
        else err = run2<2, EPI, 128, false>(ntm, 1, n256 * 256, st, A0, A1, nullptr, lda, B16, ldb, Kd, e);
    }
    return err;
}

bool logl_fuses_softmax(const emc_ctx *c) { return c->mfma && !c->chi2_mfma && !c->k_split && c->opt.mfma_tile == 0; }

int softmax_fused_bands(const emc_ctx *c) { return (int)((int64_t)((c->S + c->chunk - 1) / c->chunk) * (c->chunkp / BM2)); }

bool logl_skips_stores(const emc_ctx *c)
{
    // the softmax's 8-row groups must not straddle a band (chunks of a multiple of 8 rotations)
    return c->skip_logits && logl_fuses_softmax(c) && c->la_terms == 2 && c->opt.mfma_staging == 0 &&
           c->chunk % 8 == 0;
}

hipError_t launch_logl_mfma(const emc_ctx *c, int s0, int cs, hipStream_t st)
{
    EpiArgs e{};
    e.tile_rows = c->opt.tile_rows;
    e.C = c->R + (size_t)s0 * c->Ip;
    e.ldc = c->Ip;
    e.m_valid = cs;
    e.n_valid = c->I;
    e.bias = c->bias + s0;
    const dim3 grid(c->Ip / BN, c->chunkp / BM);
    const bf16_t *A1 = c->la_terms == 3 ? c->Lm : c->Ll;
    const bf16_t *A2 = c->Ll;
    if (logl_fuses_softmax(c)) {
        // band partials in c->part: pm (floats) [bands][I] then ps (doubles) [bands][I]
        const int64_t nb = softmax_fused_bands(c);
        e.pm = reinterpret_cast<float *>(c->part);
        e.ps = c->part + nb * c->I;
        e.band0 = (int64_t)(s0 / c->chunk) * (c->chunkp / BM2);
        e.ldp = c->I;
        if (logl_skips_stores(c)) {
            // the running maxima start at -inf with the pass's first chunk and carry over the others
            e.runmax = c->runmax;
            e.skipped = c->scal + SC_SKIPPED;
            e.blk = c->blk;
            if (s0 == 0) {
                hipError_t err = launch_fill_i32(c->runmax, kOrdNegInf, c->I, st);
                if (err == hipSuccess) err = launch_fill_f64(e.skipped, 0.0, 1, st);
                if (err != hipSuccess) return err;
            }
        }
        if (c->la_terms == 2 && c->opt.mfma_staging == 0)
            return run_dma<EPI_LOGL>(c->chunkp, c->Ip, st, c->Lh, A1, c->Pp, c->K16, c->K8, c->Pp, c->Pp, e);
        return run_large<EPI_LOGL>(c->chunkp, c->Ip, st, c->la_terms, c->Lh, A1, A2, c->Pp, c->K16, c->K8, c->Pp,
                                   c->Pp, e);
    }
    if (c->la_terms == 3) {
        if (c->k_split) return run<3, 2, EPI_LOGL>(grid, st, c->Lh, A1, A2, c->Pp, c->K16, c->K16lo, c->Pp, c->Pp, e);
        return run<3, 1, EPI_LOGL>(grid, st, c->Lh, A1, A2, c->Pp, c->K16, nullptr, c->Pp, c->Pp, e);
    }
    if (c->k_split) return run<2, 2, EPI_LOGL>(grid, st, c->Lh, A1, nullptr, c->Pp, c->K16, c->K16lo, c->Pp, c->Pp, e);
    return run<2, 1, EPI_LOGL>(grid, st, c->Lh, A1, nullptr, c->Pp, c->K16, nullptr, c->Pp, c->Pp, e);
}

hipError_t launch_gemm_store_mfma(const emc_ctx *c, int na, const bf16_t *A0, const bf16_t *A1, const bf16_t *A2,
                                  const bf16_t *B16, const uint8_t *B8, float *C, bool accumulate, hipStream_t st)
{
    if (na < 1 || na > 3 || c->chunkp % BM2 || c->Ip % 128 || c->Pp % BK2) return hipErrorInvalidValue;
    EpiArgs e{};
    e.tile_rows = c->opt.tile_rows;
    e.C = C;
    e.ldc = c->Ip;
    e.m_valid = c->chunkp;
    e.n_valid = c->Ip;
    e.accumulate = accumulate;
    if (na == 2 && B16 && c->opt.mfma_staging == 0)
        return run_dma<EPI_STORE>(c->chunkp, c->Ip, st, A0, A1, c->Pp, B16, B8, c->Pp, c->Pp, e);
    return run_large<EPI_STORE>(c->chunkp, c->Ip, st, na, A0, A1, A2, c->Pp, B16, B8, c->Pp, c->Pp, e);
}

hipError_t launch_mstep_mfma(const emc_ctx *c, int s0, int cs, bool fused, const double *gate, double gate_thr,
                             hipStream_t st)
{
    EpiArgs e{};
    e.tile_rows = c->opt.tile_rows;
    e.gate = fused ? nullptr : gate;
    e.gate_thr = gate_thr;
    e.C = c->U;
    e.ldc = c->Pp;
    e.m_valid = fused ? cs : c->chunkp;
    e.n_valid = c->Pp;
    e.T = c->T + s0;
    e.rotm = c->rotm + 12 * (size_t)s0;
    e.axyz = c->axyz;
    e.Pp = c->Pp;
    e.Pa = c->Pa;
    e.L = c->L;
    e.interp = c->opt.interp;
    e.Mnum = c->Mnum;
    e.Wt = c->Wt;
    const dim3 grid(c->Pp / BN, c->chunkp / BM);
    const bf16_t *Ah = c->Ph + (size_t)s0 * c->Ip, *Al = c->Pl + (size_t)s0 * c->Ip;
    if (fused) {
        if (c->k_split) return run<2, 2, EPI_SCATTER>(grid, st, Ah, Al, nullptr, c->Ip, c->KT16, c->KT16lo, c->Ip, c->Ip, e);
        return run<2, 1, EPI_SCATTER>(grid, st, Ah, Al, nullptr, c->Ip, c->KT16, nullptr, c->Ip, c->Ip, e);
    }
    if (c->k_split) return run<2, 2, EPI_STORE>(grid, st, Ah, Al, nullptr, c->Ip, c->KT16, c->KT16lo, c->Ip, c->Ip, e);
    if (c->opt.mfma_tile == 0 && c->opt.mfma_staging == 0)
        return run_dma<EPI_STORE>(c->chunkp, c->Pp, st, Ah, Al, c->Ip, c->KT16, c->KT8, c->Ip, c->Ip, e);
    if (c->opt.mfma_tile == 0)
        return run_large<EPI_STORE>(c->chunkp, c->Pp, st, 2, Ah, Al, nullptr, c->Ip, c->KT16, c->KT8, c->Ip, c->Ip, e);
    return run<2, 1, EPI_STORE>(grid, st, Ah, Al, nullptr, c->Ip, c->KT16, nullptr, c->Ip, c->Ip, e);
}

}  // namespace emc
