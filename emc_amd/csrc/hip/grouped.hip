// Grouped contractions of a proposal iteration: each pattern is scored only against the fine
// rotations of its selected coarse cells.
//
// No reference counterpart (main.cpp:239-480 scores every rotation for every pattern); each
// group is the dense stage restricted to one cell, so the arithmetic per (rotation, pattern)
// pair is the dense MFMA path's (same operands, same k order, same epilogue).
//
// Work lists (built on the host per iteration): cell c owns the fine rotations
// cell_rot[rbeg[c] .. rbeg[c+1]) (nr_c of them) and the "slots" sbeg[c] .. sbeg[c+1) (ns_c
// patterns that selected c; slot_pat gives the pattern).  Per cell, logits and responsibilities
// are a row-major block [nr_c][nsp_c] (nsp_c = ns_c rounded up to 32, padding columns hold P = 0)
// at element offset boff[c] of c->R (logits) and c->Ph / c->Pl (P hi / lo).
//
//   logL   : block(c)[r][n] = sum_j logW~[rot(r), j] K[pat(n), j] + bias[rot(r)]    (rows gathered)
//   M-step : U[rot(r), j]   = sum_n P(c)[r][n] K[pat(n), j]      (count rows gathered, transposed
//                                                                  into LDS while staged)
// Tiles are 128 x 128, 256 threads = 4 wavefronts each 64 x 64 (2 x 2 v_mfma_f32_32x32x16_bf16),
// BK = 32 with two LDS stages (rows padded to 40 BF16), workgroups renumbered per XCD.
#include <algorithm>

#include "../emc_internal.h"
#include "sparse_row.h"

namespace emc {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

constexpr int TM = 128, TN = 128, TK = 32, TLD = TK + 8;

struct GTab {
    const int32_t *cell_rot, *rbeg, *cell_of;           // partition of the fine rotations
    const int32_t *slot_pat, *sbeg, *nsp;                // slots per cell
    const int32_t *pbeg, *pslot;                         // slots per pattern
    const int64_t *boff;                                 // block offsets
    const int4 *tiles;                                   // (cell, m0, n0, -)
};

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

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

enum { G_LOGL = 0, G_MSTEP = 1 };

// MODE G_LOGL : A = NA log-slice terms [S][lda] (rows gathered by cell_rot), B = counts [I][ldb]
//               (rows gathered by slot_pat), k extent Kd = Pp.
// MODE G_MSTEP: A = P hi, lo blocks (NA = 2), B = counts [I][ldb] read as B(j, n) = K[pat(n)][j]
//               (each thread stages 8 pixels of 2 slots and writes them as slot pairs), k extent
//               nsp_c.
template <int MODE, int NA, bool BU8>
__global__ __launch_bounds__(256) void k_grouped(const bf16_t *__restrict__ A0, const bf16_t *__restrict__ A1,
                                                 const bf16_t *__restrict__ A2, int64_t lda,
                                                 const void *__restrict__ Bv, int64_t ldb, int Pp, GTab g,
                                                 const float *__restrict__ bias, float *__restrict__ out,
                                                 const double *__restrict__ gate, double gate_thr)
{
    if (MODE == G_MSTEP && gate && *gate <= gate_thr) return;  // the sparse M-step runs instead
    constexpr int STAGE = (NA * TM + TN) * TLD;
    __shared__ __attribute__((aligned(16))) bf16_t smem[2 * STAGE];
    __shared__ float sbias[TM];
    __shared__ int srot[TM];

    const int nwg = gridDim.x, bid = blockIdx.x;
    const int base = nwg / 8, rem = nwg % 8, xcd = bid % 8, loc = bid / 8;
    const int t = xcd * base + min(xcd, rem) + loc;
    const int4 td = g.tiles[t];
    const int c = td.x, m0 = td.y, n0 = td.z;
    const int rb = g.rbeg[c], nr = g.rbeg[c + 1] - rb;
    const int sb = g.sbeg[c], ns = g.sbeg[c + 1] - sb, nsp = g.nsp[c];
    const int64_t boff = g.boff[c];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int lrow = tid >> 2, lcol = (tid & 3) * 8;

    if (tid < TM) {
        const int rot = g.cell_rot[rb + min(m0 + tid, nr - 1)];
        srot[tid] = rot;
        if (MODE == G_LOGL) sbias[tid] = bias[rot];
    }

    const bf16_t *ga[NA][2];
    const unsigned char *gb[2];
    int Kd;
    if (MODE == G_LOGL) {
        Kd = Pp;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int64_t arow = g.cell_rot[rb + min(m0 + lrow + 64 * h, nr - 1)];
            ga[0][h] = A0 + arow * lda + lcol;
            if (NA > 1) ga[1][h] = A1 + arow * lda + lcol;
            if (NA > 2) ga[NA - 1][h] = A2 + arow * lda + lcol;
            const int64_t brow = g.slot_pat[sb + min(n0 + lrow + 64 * h, ns - 1)];
            gb[h] = reinterpret_cast<const unsigned char *>(Bv) + (brow * ldb + lcol) * (BU8 ? 1 : 2);
        }
    } else {
        Kd = nsp;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int64_t arow = boff + (int64_t)min(m0 + lrow + 64 * h, nr - 1) * nsp + lcol;
            ga[0][h] = A0 + arow;
            if (NA > 1) ga[1][h] = A1 + arow;
        }
    }
    // M-step B staging: slots 2 kq, 2 kq + 1 of the k block, pixels n0 + 8 ch .. + 7; padding
    // slots read the cell's last pattern (their P is 0)
    const int kq = tid & 15, ch = tid >> 4;
    const int32_t *spat = g.slot_pat + sb;
    const unsigned char *kb = reinterpret_cast<const unsigned char *>(Bv) + (int64_t)(n0 + 8 * ch) * (BU8 ? 1 : 2);

    u32x4 ra[NA][2], rb16[2];
    u32x2 rb8[2];
    auto gload = [&](int k0) {
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int h = 0; h < 2; ++h) ra[a][h] = *reinterpret_cast<const u32x4 *>(ga[a][h] + k0);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (MODE == G_LOGL) {
                if (BU8) rb8[h] = *reinterpret_cast<const u32x2 *>(gb[h] + k0);
                else rb16[h] = *reinterpret_cast<const u32x4 *>(gb[h] + (size_t)k0 * 2);
            } else {
                const int64_t row = spat[min(k0 + 2 * kq + h, ns - 1)];
                const unsigned char *src = kb + row * ldb * (BU8 ? 1 : 2);
                if (BU8) rb8[h] = *reinterpret_cast<const u32x2 *>(src);
                else rb16[h] = *reinterpret_cast<const u32x4 *>(src);
            }
        }
    };
    auto sstore = [&](int stage) {
        bf16_t *S = smem + stage * STAGE;
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int h = 0; h < 2; ++h)
                *reinterpret_cast<u32x4 *>(S + ((size_t)a * TM + lrow + 64 * h) * TLD + lcol) = ra[a][h];
        if (MODE == G_LOGL) {
#pragma unroll
            for (int h = 0; h < 2; ++h)
                *reinterpret_cast<u32x4 *>(S + ((size_t)NA * TM + lrow + 64 * h) * TLD + lcol) =
                    BU8 ? widen_u8x8(rb8[h]) : rb16[h];
        } else {
            const u32x4 w0 = BU8 ? widen_u8x8(rb8[0]) : rb16[0], w1 = BU8 ? widen_u8x8(rb8[1]) : rb16[1];
#pragma unroll
            for (int q = 0; q < 8; ++q) {  // pixel 8 ch + q: (slot 2 kq, slot 2 kq + 1) as one word
                const uint32_t lo = (w0[q >> 1] >> (16 * (q & 1))) & 0xffffu;
                const uint32_t hi = (w1[q >> 1] >> (16 * (q & 1))) & 0xffffu;
                *reinterpret_cast<uint32_t *>(S + ((size_t)NA * TM + 8 * ch + q) * TLD + 2 * kq) = lo | (hi << 16);
            }
        }
    };

    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int nk = Kd / TK;
    const int fr = lane & 31, fk = (lane >> 5) * 8;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * TK);
        const bf16_t *S = smem + cur * STAGE;
#pragma unroll
        for (int kk = 0; kk < TK; kk += 16) {
            bf16x8 fa[NA][2], fb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int a = 0; a < NA; ++a)
                    fa[a][i] = *reinterpret_cast<const bf16x8 *>(S + ((size_t)a * TM + wm + i * 32 + fr) * TLD + kk + fk);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                fb[j] = *reinterpret_cast<const bf16x8 *>(S + ((size_t)NA * TM + wn + j * 32 + fr) * TLD + kk + fk);
#pragma unroll
            for (int a = NA - 1; a >= 0; --a)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[a][i], fb[j], acc[i][j], 0, 0, 0);
        }
        if (kt + 1 < nk) sstore(cur ^ 1);
        __syncthreads();
    }

    // accumulator: lane holds column lane % 32 and rows 8*(r/4) + 4*(lane/32) + r%4
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int ml = wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (m0 + ml >= nr) continue;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int n = n0 + wn + j * 32 + (lane & 31);
                if (MODE == G_LOGL) {
                    if (n < ns) out[boff + (int64_t)(m0 + ml) * nsp + n] = acc[i][j][r] + sbias[ml];
                } else {
                    out[(int64_t)srot[ml] * Pp + n] = acc[i][j][r];
                }
            }
        }
}

// Softmax partials: one wavefront per 64 slots of a cell (tiles (cell, s0)); online (max, FP64
// sum) over the cell's rotation rows, as the dense partial pass does over a band.
__global__ __launch_bounds__(256) void k_gsm_partial(const float *__restrict__ Lb, GTab g, int n_tiles,
                                                     float *__restrict__ pm, double *__restrict__ ps)
{
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (t >= n_tiles) return;
    const int4 td = g.tiles[t];
    const int c = td.x, s = td.y + lane;
    const int sb = g.sbeg[c], ns = g.sbeg[c + 1] - sb, nsp = g.nsp[c];
    const int nr = g.rbeg[c + 1] - g.rbeg[c];
    if (s >= ns) return;
    const float *col = Lb + g.boff[c] + s;
    float m = -1.0e10f;
    double sum = 0.0;
    for (int r = 0; r < nr; ++r) {
        const float v = col[(int64_t)r * nsp];
        if (v > m) {
            sum = sum * (double)expf(m - v) + 1.0;
            m = v;
        } else {
            const float d = v - m;
            if (d > -1.0e10f) sum += (double)expf(d);
        }
    }
    pm[sb + s] = m;
    ps[sb + s] = sum;
}

// Per pattern: combine its slots' partials in FP64 -> lse'_i; per-block sums of lse'.
__global__ __launch_bounds__(256) void k_gsm_final(const float *__restrict__ pm, const double *__restrict__ ps, GTab g,
                                                   int I, double *__restrict__ lse, double *__restrict__ blk_sum,
                                                   double *__restrict__ scal)
{
    __shared__ double red[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double l = 0.0;
    if (i < I) {
        const int q0 = g.pbeg[i], q1 = g.pbeg[i + 1];
        float m = -1.0e10f;
        for (int q = q0; q < q1; ++q) m = fmaxf(m, pm[g.pslot[q]]);
        double sum = 0.0;
        for (int q = q0; q < q1; ++q) {
            const int sl = g.pslot[q];
            const double sv = ps[sl];
            if (sv > 0.0) sum += sv * exp((double)pm[sl] - (double)m);
        }
        l = (double)m + log(sum);
        if (!isfinite(l)) scal[SC_NONFINITE] = 1.0;
        lse[i] = l;
    }
    l = wave_sum(l);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) blk_sum[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void k_slot_lse(const int32_t *__restrict__ slot_pat, const double *__restrict__ lse, int64_t n,
                           double *__restrict__ slse)
{
    const int64_t s = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (s < n) slse[s] = lse[slot_pat[s]];
}

// P = exp(logit - lse) as BF16 hi + lo in the block layout (0 in padding columns) and the slice
// sums T[rot] (FP64).  One wavefront per fine rotation (in cell order); rotations of cells no
// pattern selected get T = 0, which the insertion skips.
__global__ __launch_bounds__(256) void k_gsm_apply(const float *__restrict__ Lb, GTab g, int S,
                                                   const double *__restrict__ slse, bf16_t *__restrict__ Ph,
                                                   bf16_t *__restrict__ Pl, double *__restrict__ T,
                                                   double *__restrict__ nzrow)
{
    const int gi = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (gi >= S) return;
    const int rot = g.cell_rot[gi], c = g.cell_of[rot], r = gi - g.rbeg[c];
    const int sb = g.sbeg[c], ns = g.sbeg[c + 1] - sb, nsp = g.nsp[c];
    if (ns == 0) {
        if (lane == 0) {
            T[rot] = 0.0;
            nzrow[gi] = 0.0;
        }
        return;
    }
    const int64_t row = g.boff[c] + (int64_t)r * nsp;
    double acc = 0.0;
    int nz = 0;  // non-zero P of the row (for the choice of the M-step form)
    for (int s = lane; s < nsp; s += 64) {
        float p = 0.0f;
        if (s < ns) {  // the dense softmax's arithmetic: lse as a float pair, exp on the hardware exp2
            const double l = slse[sb + s];
            const float lh = (float)l, ll = (float)(l - (double)lh);
            p = __builtin_amdgcn_exp2f(((Lb[row + s] - lh) - ll) * 1.4426950408889634f);
        }
        const bf16_t h = f2bf_dev(p), lo = f2bf_dev(p - bf2f(h));
        Ph[row + s] = h;
        Pl[row + s] = lo;
        acc += (double)p;
        nz += (h | lo) != 0;
    }
    acc = wave_sum(acc);
    nz = wave_sum(nz);
    if (lane == 0) {
        T[rot] = acc;
        nzrow[gi] = (double)nz;
    }
}

// Sparse M-step of a proposal iteration: the U row of each fine rotation of a selected cell from
// the non-zero responsibilities of its row of the cell's P block (sparse_row.h; entry n of the row
// is the pattern slot_pat[sbeg[c] + n]).  One workgroup per fine rotation, in cell order; rotations
// of cells no pattern selected are left alone (their T is 0 and the insertion skips them).
template <int KT, int NPASS>
__global__ __launch_bounds__(256) void k_gmstep_sparse(const bf16_t *__restrict__ Ph, const bf16_t *__restrict__ Pl,
                                                       GTab g, const void *__restrict__ Kv,
                                                       const bf16_t *__restrict__ Klo, int Pp,
                                                       const double *__restrict__ gate, double gate_thr,
                                                       float *__restrict__ U)
{
    if (gate && *gate > gate_thr) return;  // the dense grouped contraction runs instead
    const int gi = blockIdx.x;
    const int rot = g.cell_rot[gi], c = g.cell_of[rot], r = gi - g.rbeg[c];
    const int sb = g.sbeg[c], ns = g.sbeg[c + 1] - sb, nsp = g.nsp[c];
    if (ns == 0) return;
    const int64_t row = g.boff[c] + (int64_t)r * nsp;
    sparse_row<KT, NPASS>(Ph + row, Pl + row, ns, g.slot_pat + sb, Kv, Klo, Pp, U + (size_t)rot * Pp);
}

// The kMaxRecall cells of largest score seen so far, sorted by score (descending; on equal
// scores the cell offered first stays first).  Static indexing only, so the arrays stay in
// registers.
struct TopCells {
    float v[kMaxRecall];
    int c[kMaxRecall];
    __device__ void init()
    {
#pragma unroll
        for (int q = 0; q < kMaxRecall; ++q) {
            v[q] = -INFINITY;
            c[q] = -1;
        }
    }
    __device__ void offer(float x, int cell)
    {
#pragma unroll
        for (int q = 0; q < kMaxRecall; ++q) {
            if (x > v[q]) {
                const float tv = v[q];
                const int tc = c[q];
                v[q] = x;
                c[q] = cell;
                x = tv;
                cell = tc;
            }
        }
    }
    __device__ void store(int32_t *out, int R) const
    {
#pragma unroll
        for (int q = 0; q < kMaxRecall; ++q)
            if (q < R) out[q] = c[q];
    }
};

// Recall after a full search: the cells of largest max logit, Mx [C][ld] (-inf for an empty
// cell, never recalled).  One thread per pattern (reads coalesced across patterns).
__global__ __launch_bounds__(256) void k_recall_dense(const float *__restrict__ Mx, int64_t ld, int I, int C, int R,
                                                      int32_t *__restrict__ out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= I) return;
    TopCells t;
    t.init();
    for (int c = 0; c < C; ++c) t.offer(Mx[(size_t)c * ld + i], c);
    t.store(out + (size_t)i * R, R);
}

// Recall after a proposal iteration: the searched cells of largest max logit (the slot partial
// maximum pm).  The cell of a slot is found by binary search in sbeg.
__global__ __launch_bounds__(256) void k_recall_slots(const float *__restrict__ pm, GTab g, int I, int C, int R,
                                                      int32_t *__restrict__ out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= I) return;
    TopCells t;
    t.init();
    for (int q = g.pbeg[i]; q < g.pbeg[i + 1]; ++q) {
        const int sl = g.pslot[q];
        int lo = 0, hi = C;  // largest c with sbeg[c] <= sl
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (g.sbeg[mid] <= sl) lo = mid;
            else hi = mid;
        }
        t.offer(pm[sl], lo);
    }
    t.store(out + (size_t)i * R, R);
}

// Doubles at the start of c->part used by the grouped softmax's per-block lse sums.
int64_t blk_pad(const emc_ctx *c) { return ((c->I + 255) / 256 + 16) / 16 * 16; }

GTab tables(const emc_ctx *c, const int4 *tiles)
{
    const Proposal *p = c->prop;
    GTab g;
    g.cell_rot = p->cell_rot;
    g.rbeg = p->rbeg;
    g.cell_of = p->cell_of;
    g.slot_pat = p->lists + p->off_slot_pat;
    g.sbeg = p->lists + p->off_sbeg;
    g.nsp = p->lists + p->off_nsp;
    g.pbeg = p->lists + p->off_pbeg;
    g.pslot = p->lists + p->off_pslot;
    g.boff = p->boff;
    g.tiles = tiles;
    return g;
}

const int4 *tile_list(const Proposal *p, int64_t off) { return reinterpret_cast<const int4 *>(p->lists + off); }

}  // namespace

hipError_t launch_grouped_logl(const emc_ctx *c, hipStream_t st)
{
    const Proposal *p = c->prop;
    if (p->n_tiles_e == 0) return hipSuccess;
    const GTab g = tables(c, tile_list(p, p->off_tile_e));
    const dim3 grid(p->n_tiles_e), blk(256);
    const int64_t ld = c->Pp;
    const bf16_t *A1 = c->la_terms == 3 ? c->Lm : c->Ll;
    if (c->K8) {
        if (c->la_terms == 3)
            hipLaunchKernelGGL((k_grouped<G_LOGL, 3, true>), grid, blk, 0, st, c->Lh, A1, c->Ll, ld, c->K8, ld, c->Pp, g,
                               c->bias, c->R, nullptr, 0.0);
        else
            hipLaunchKernelGGL((k_grouped<G_LOGL, 2, true>), grid, blk, 0, st, c->Lh, A1, nullptr, ld, c->K8, ld, c->Pp,
                               g, c->bias, c->R, nullptr, 0.0);
    } else {
        if (c->la_terms == 3)
            hipLaunchKernelGGL((k_grouped<G_LOGL, 3, false>), grid, blk, 0, st, c->Lh, A1, c->Ll, ld, c->K16, ld, c->Pp,
                               g, c->bias, c->R, nullptr, 0.0);
        else
            hipLaunchKernelGGL((k_grouped<G_LOGL, 2, false>), grid, blk, 0, st, c->Lh, A1, nullptr, ld, c->K16, ld,
                               c->Pp, g, c->bias, c->R, nullptr, 0.0);
    }
    return hipGetLastError();
}

hipError_t launch_grouped_softmax(const emc_ctx *c, hipStream_t st)
{
    const Proposal *p = c->prop;
    GTab g = tables(c, tile_list(p, p->off_tile_sm));
    hipLaunchKernelGGL(k_gsm_partial, dim3((p->n_tiles_sm + 3) / 4), dim3(256), 0, st, c->R, g, p->n_tiles_sm,
                       p->slot_pm, p->slot_ps);
    const int ncol = (c->I + 255) / 256;
    double *blk = c->part;
    hipLaunchKernelGGL(k_gsm_final, dim3(ncol), dim3(256), 0, st, p->slot_pm, p->slot_ps, g, c->I, c->lse, blk, c->scal);
    hipError_t e = launch_sum_f64(blk, ncol, c->scal + SC_TOTAL, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_slot_lse, dim3((unsigned)((p->n_slots + 255) / 256)), dim3(256), 0, st, g.slot_pat, c->lse,
                       p->n_slots, p->slot_lse);
    double *nzrow = c->part + blk_pad(c);  // one non-zero count per fine rotation
    hipLaunchKernelGGL(k_gsm_apply, dim3((c->S + 3) / 4), dim3(256), 0, st, c->R, g, c->S, p->slot_lse, c->Ph, c->Pl,
                       c->T, nzrow);
    e = launch_mark_p_dirty(c, st);  // Ph / Pl now hold the grouped layout
    if (e != hipSuccess) return e;
    return launch_sum_f64(nzrow, c->S, c->scal + SC_NNZ, st);
}

hipError_t launch_recall_grouped(const emc_ctx *c, hipStream_t st)
{
    const Proposal *p = c->prop;
    const int R = p->o.recall;
    if (R < 1 || R > kMaxRecall) return R == 0 ? hipSuccess : hipErrorInvalidValue;
    const GTab g = tables(c, nullptr);
    hipLaunchKernelGGL(k_recall_slots, dim3((c->I + 255) / 256), dim3(256), 0, st, p->slot_pm, g, c->I, p->C, R,
                       p->recall);
    return hipGetLastError();
}

hipError_t launch_recall_dense(const emc_ctx *c, hipStream_t st)
{
    const Proposal *p = c->prop;
    const int R = p->o.recall;
    if (R < 1 || R > kMaxRecall) return R == 0 ? hipSuccess : hipErrorInvalidValue;
    hipLaunchKernelGGL(k_recall_dense, dim3((c->I + 255) / 256), dim3(256), 0, st, p->Mx, (int64_t)c->Ip, c->I, p->C, R,
                       p->recall);
    return hipGetLastError();
}

hipError_t launch_grouped_mstep(const emc_ctx *c, hipStream_t st)
{
    const Proposal *p = c->prop;
    if (p->n_tiles_m == 0) return hipSuccess;
    // dense or sparse form, chosen on the device from the non-zero count of the grouped softmax
    // (at most 2 % of the block elements non-zero: sparse), unless the options fix one
    const int mode = c->Pp > kSparseMaxPp ? EMC_MSTEP_DENSE : c->opt.mstep;
    const double *gate = mode == EMC_MSTEP_AUTO ? c->scal + SC_NNZ : nullptr;
    const double thr = 0.02 * (double)p->n_blk;
    if (mode != EMC_MSTEP_DENSE) {
        const GTab gs = tables(c, nullptr);
        const bool np1 = c->Pp <= 4096;
        constexpr int NP = kSparseMaxPp / 4096;
        if (c->K8) {
            if (np1)
                hipLaunchKernelGGL((k_gmstep_sparse<0, 1>), dim3(c->S), dim3(256), 0, st, c->Ph, c->Pl, gs, c->K8, nullptr,
                                   c->Pp, gate, thr, c->U);
            else
                hipLaunchKernelGGL((k_gmstep_sparse<0, NP>), dim3(c->S), dim3(256), 0, st, c->Ph, c->Pl, gs, c->K8, nullptr,
                                   c->Pp, gate, thr, c->U);
        } else {
            if (np1)
                hipLaunchKernelGGL((k_gmstep_sparse<1, 1>), dim3(c->S), dim3(256), 0, st, c->Ph, c->Pl, gs, c->K16, nullptr,
                                   c->Pp, gate, thr, c->U);
            else
                hipLaunchKernelGGL((k_gmstep_sparse<1, NP>), dim3(c->S), dim3(256), 0, st, c->Ph, c->Pl, gs, c->K16, nullptr,
                                   c->Pp, gate, thr, c->U);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (mode == EMC_MSTEP_SPARSE) return hipSuccess;
    const GTab g = tables(c, tile_list(p, p->off_tile_m));
    const dim3 grid(p->n_tiles_m), blk(256);
    if (c->K8)
        hipLaunchKernelGGL((k_grouped<G_MSTEP, 2, true>), grid, blk, 0, st, c->Ph, c->Pl, nullptr, 0, c->K8, c->Pp, c->Pp,
                           g, nullptr, c->U, gate, thr);
    else
        hipLaunchKernelGGL((k_grouped<G_MSTEP, 2, false>), grid, blk, 0, st, c->Ph, c->Pl, nullptr, 0, c->K16, c->Pp, c->Pp,
                           g, nullptr, c->U, gate, thr);
    return hipGetLastError();
}

}  // namespace emc
