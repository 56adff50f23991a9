// Maximize, scatter half, binned: deterministic insertion without atomics.
//
// The geometry of insertion (which voxels a rotated pixel touches, with which trilinear weights)
// depends only on the rotation list, the detector coordinates and the mask, not on the
// iteration.  Once per geometry the valid (rotation s, active pixel j) pairs are sorted by the
// model brick (kBrick^3 voxels) that holds their base voxel, per rotation chunk, with a stable
// radix sort (rocPRIM), so each brick lists its pairs in (s, j) order; each brick's list is cut
// into units of at most kUnit pairs, and inside a unit the pairs are ranked once more by base
// voxel (a 16-bit permutation, the start of each base voxel's run, and for trilinear the pairs'
// fractions in base-voxel order, laid out for coalesced reads).  Per pair the iteration then reads
// 4 bytes of list (the packed pair), 2 of permutation and 12 of fractions.  Every iteration:
//   1. one workgroup per unit reads its pairs in (s, j) order, where the reads of U are runs of
//      consecutive pixels, and writes U and T_s into LDS at the pair's position in base-voxel order;
//   2. every thread sums an equal share of the positions into per-base-voxel corner sums (the
//      pieces of a voxel's run that crosses threads are added in thread order);
//   3. a workgroup runs steps 1-2 for up to kUnitsPerBlock units of one brick in turn, then every
//      voxel of the halo brick ((kBrick+1)^3) gathers the corner sums in a fixed order, and the
//      halo brick is stored as the block's partial (no atomics);
//   4. a grouping pass sums the partials of up to kGroup consecutive blocks of one brick, and a
//      last pass gives every model voxel the sum, in order, of the groups of the (at most 8)
//      bricks whose halo covers it.
// No sum depends on scheduling, so repeated runs give bit-identical models.  This replaces the 16
// global atomics per (s, j) of the reference's insert_slices_kernel (emc_atomic.cu:33-85,
// 183-200); the skip rules (T_s > 1e-10, emc_atomic.cu:190; U >= 0, emc_atomic.cu:56) and the
// bounds are the same as in the atomic form.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "../emc_internal.h"
#include "geom.h"

namespace emc {
namespace {

constexpr int kBrick = 8;
constexpr int kBrick3 = kBrick * kBrick * kBrick;  // base voxels of a brick
constexpr int kUnit = 4096;                        // pairs per workgroup (positions fit 16 bits)
constexpr int kVbeg = 520;                         // run starts per unit: kBrick3 + 1, padded to 16 bytes
constexpr int kUnitsPerBlock = 4;                  // units one insertion workgroup processes in turn
constexpr int kGroup = 32;                         // blocks whose partials one grouping workgroup sums
constexpr int kHalo = kBrick + 1;
constexpr int kHalo3 = kHalo * kHalo * kHalo;      // voxels a brick's pairs reach
constexpr int kStage = 8;                          // pairs per thread whose loads are issued together
constexpr int kThreads = 512;  // 8 waves; with 67 KB of LDS two workgroups (16 waves) fit a CU
static_assert(kUnit % kThreads == 0, "step 2 gives every thread the same share");

struct BinGeom {
    int L, nb, S, Pa, Pp, chunk, interp, jbits;
};

__device__ __forceinline__ void unpack(uint32_t p, int jbits, int &s, int &j)
{
    s = (int)(p >> jbits);
    j = (int)(p & ((1u << jbits) - 1u));
}

// Voxel cell of active pixel j under rotation s (nearest: the voxel; trilinear: the base corner and
// the fractions); false when out of bounds (emc_atomic.cu:67-69 for nearest).
__device__ __forceinline__ bool item_cell(const float *__restrict__ rotm, const float *__restrict__ axyz,
                                          const BinGeom &g, int s, int j, int &bx, int &by, int &bz, float &fx,
                                          float &fy, float &fz)
{
    float m[9];
    load_rot(rotm, s, m);
    float ux, uy, uz;
    rot_point(m, axyz[j], axyz[g.Pp + j], axyz[2 * g.Pp + j], model_c(g.L), ux, uy, uz);
    if (g.interp == EMC_INTERP_NEAREST) {
        bx = nearest(ux);
        by = nearest(uy);
        bz = nearest(uz);
        fx = fy = fz = 0.0f;
        return bx >= 0 && bx < g.L && by >= 0 && by < g.L && bz >= 0 && bz < g.L;
    }
    return tri_axis(ux, g.L, bx, fx) && tri_axis(uy, g.L, by, fy) && tri_axis(uz, g.L, bz, fz);
}

// Sort key of every pair (chunk-major brick; invalid pairs get invalid_key) and its packed value
// s << jbits | j, in (s, j) order.
__global__ void k_bin_keys(const float *__restrict__ rotm, const float *__restrict__ axyz, BinGeom g, int64_t n,
                           uint32_t invalid_key, uint32_t *__restrict__ keys, uint32_t *__restrict__ vals)
{
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * blockDim.x) {
        const int s = (int)(idx / g.Pa), j = (int)(idx % g.Pa);
        int bx, by, bz;
        float fx, fy, fz;
        uint32_t key = invalid_key;
        if (item_cell(rotm, axyz, g, s, j, bx, by, bz, fx, fy, fz)) {
            const int brick = ((bz / kBrick) * g.nb + by / kBrick) * g.nb + bx / kBrick;
            key = (uint32_t)((s / g.chunk) * g.nb * g.nb * g.nb + brick);
        }
        keys[idx] = key;
        vals[idx] = ((uint32_t)s << g.jbits) | (uint32_t)j;
    }
}

__global__ void k_bin_count(const uint32_t *__restrict__ keys, int64_t n, uint32_t nkeys, int32_t *__restrict__ counts)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t k = keys[t];
        if (k < nkeys && (t == 0 || keys[t - 1] != k)) counts[2 * k] = (int32_t)t;              // first position
        if (k < nkeys && (t == n - 1 || keys[t + 1] != k)) counts[2 * k + 1] = (int32_t)(t + 1);  // end
    }
}

__global__ void k_iota(uint32_t *__restrict__ v, int64_t n)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        v[t] = (uint32_t)t;
}

// Second sort key of the pairs of unit u: u * kBrick3 + the base voxel's index inside the brick.
__global__ __launch_bounds__(kThreads) void k_unit_keys(const int32_t *__restrict__ units,
                                                        const uint32_t *__restrict__ vals,
                                                        const float *__restrict__ rotm,
                                                        const float *__restrict__ axyz, BinGeom g,
                                                        uint32_t *__restrict__ keys)
{
    const int u = blockIdx.x;
    const int begin = units[4 * u + 1], end = units[4 * u + 2];
    for (int t = begin + threadIdx.x; t < end; t += kThreads) {
        int s, j;
        unpack(vals[t], g.jbits, s, j);
        int bx = 0, by = 0, bz = 0;
        float fx, fy, fz;
        item_cell(rotm, axyz, g, s, j, bx, by, bz, fx, fy, fz);  // valid: only valid pairs are listed
        const int lv = ((bz % kBrick) * kBrick + by % kBrick) * kBrick + bx % kBrick;
        keys[t] = (uint32_t)u * kBrick3 + (uint32_t)lv;
    }
}

// Per unit: each pair's position in base-voxel order (a stable order, so (s, j) order within a base
// voxel), stored at the pair's place in (s, j) order; the start of each base voxel's run
// (kBrick3 + 1 entries); for trilinear, the fractions of the pair at base-voxel position k, stored
// at frac[(u * 3 + axis) * kUnit + (k % kPer) * kThreads + k / kPer], where the insertion's thread
// k / kPer reads them in step (k % kPer) of its share, all lanes of a wave on consecutive words.
__global__ __launch_bounds__(kThreads) void k_unit_perm(const int32_t *__restrict__ units,
                                                        const uint32_t *__restrict__ skeys,
                                                        const uint32_t *__restrict__ spos,
                                                        const uint32_t *__restrict__ vals,
                                                        const float *__restrict__ rotm,
                                                        const float *__restrict__ axyz, BinGeom g,
                                                        uint16_t *__restrict__ inv, uint16_t *__restrict__ vbeg,
                                                        float *__restrict__ frac)
{
    constexpr int kPer = kUnit / kThreads;
    const int u = blockIdx.x;
    const int begin = units[4 * u + 1], end = units[4 * u + 2];
    for (int t = begin + threadIdx.x; t < end; t += kThreads) {
        const int k = t - begin;
        inv[spos[t]] = (uint16_t)k;
        if (frac) {
            int s, j, bx, by, bz;
            float f[3];
            unpack(vals[spos[t]], g.jbits, s, j);
            item_cell(rotm, axyz, g, s, j, bx, by, bz, f[0], f[1], f[2]);
            for (int a = 0; a < 3; ++a)
                frac[((size_t)u * 3 + a) * kUnit + (k % kPer) * kThreads + k / kPer] = f[a];
        }
    }
    for (int lv = threadIdx.x; lv <= kBrick3; lv += kThreads) {
        const uint32_t key = (uint32_t)u * kBrick3 + (uint32_t)lv;
        int lo = begin, hi = end;  // first position with skeys >= key
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (skeys[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        vbeg[(size_t)u * kVbeg + lv] = (uint16_t)(lo - begin);
    }
}

// Steps 1-3 for one block: up to kUnitsPerBlock consecutive units of one brick, one after the other,
// whose corner sums accumulate in LDS; the block's halo brick is stored as its partial
// (part: [block - first block of the chunk][2][kHalo3], numerator then weight).
// Step 1 reads a unit's pairs in (s, j) order (coalesced: runs of pixels of one rotation) and
// writes U (or -1: skipped) and T_s into LDS at the pair's position in base-voxel order.  Step 2
// gives every thread the same share of the unit, kPer consecutive positions, with the trilinear
// fractions read from the precomputed coalesced planes (a thread per base voxel would wait on the
// brick's heaviest voxel: at 64^3 / 25,680 the slowest thread of a unit would sum 2.5 times the
// mean, up to 32 times).  A base voxel's run that crosses a thread boundary is summed piecewise: a
// piece that continues into the next thread is handed over as a tail, and the thread holding the
// run's last position adds the pieces in thread order.  Corner c = dx + 2 dy + 4 dz, weights
// (1-f or f) per axis.  LDS layouts avoid bank conflicts: position k sits at k ^ ((k >> 3) & 7) (a
// thread's kPer = 8 positions are consecutive, so the 64 lanes of a read fall on 64 banks), the
// corner sums are corner-major, the tails value-major.
__device__ __forceinline__ int rswz(int k) { return k ^ ((k >> 3) & 7); }

template <int INTERP>
__global__ __launch_bounds__(kThreads, 2) void k_insert_binned(const int32_t *__restrict__ blocks,
                                                               const int32_t *__restrict__ units,
                                                               const uint32_t *__restrict__ vals,
                                                               const uint16_t *__restrict__ inv,
                                                               const uint16_t *__restrict__ vbeg,
                                                               const float *__restrict__ frac,
                                                               const float *__restrict__ U, int s0,
                                                               const double *__restrict__ T, BinGeom g,
                                                               float *__restrict__ part)
{
    constexpr int NC = INTERP == EMC_INTERP_NEAREST ? 1 : 8;
    constexpr int kPer = kUnit / kThreads;  // positions per thread in step 2
    static_assert(kPer == 8, "the position swizzle assumes 8 consecutive positions per thread");
    static_assert(2 * kUnit >= kThreads * 2 * NC, "the tails reuse the staging space");
    __shared__ float sv[2][kUnit];  // U (-1: skipped) and T_s by base-voxel position
    __shared__ uint16_t svb[kBrick3 + 1];
    __shared__ float cM[NC][kBrick3], cW[NC][kBrick3];
    __shared__ int tail_lv[kThreads];
    float(*tail)[kThreads] = reinterpret_cast<float(*)[kThreads]>(&sv[0][0]);  // [2 NC][kThreads]
    const int b = blockIdx.x;
    const int ub = blocks[4 * b + 1], ue = blocks[4 * b + 2];
    const int tid = threadIdx.x;
    for (int k = tid; k < NC * kBrick3; k += kThreads) {
        (&cM[0][0])[k] = 0.0f;
        (&cW[0][0])[k] = 0.0f;
    }
    for (int u = ub; u < ue; ++u) {
        const int begin = units[4 * u + 1], end = units[4 * u + 2];
        const int n = end - begin;
        // this thread's fractions (issued before the barrier: they do not depend on LDS)
        const int k0 = tid * kPer, k1 = min(n, k0 + kPer);
        float fr[kPer][3];
        if constexpr (INTERP == EMC_INTERP_TRILINEAR) {
#pragma unroll
            for (int i = 0; i < kPer; ++i)
#pragma unroll
                for (int a = 0; a < 3; ++a) fr[i][a] = frac[((size_t)u * 3 + a) * kUnit + i * kThreads + tid];
        }
        __syncthreads();  // the previous unit's tails and run starts are consumed
        for (int k = tid; k <= kBrick3; k += kThreads) svb[k] = vbeg[(size_t)u * kVbeg + k];
        // 1. U and T_s by base-voxel position, from the pairs in (s, j) order
        for (int q0 = tid; q0 < n; q0 += kThreads * kStage) {
            uint32_t pv[kStage];
            uint16_t pos[kStage];
#pragma unroll
            for (int q = 0; q < kStage; ++q) {
                const int k = begin + min(q0 + q * kThreads, n - 1);  // past the end: a valid pair, not stored
                pv[q] = vals[k];
                pos[q] = inv[k];
            }
            double Td[kStage];
            float val[kStage];
#pragma unroll
            for (int q = 0; q < kStage; ++q) {
                int s, j;
                unpack(pv[q], g.jbits, s, j);
                Td[q] = T[s];
                val[q] = U[(size_t)(s - s0) * g.Pp + j];
            }
#pragma unroll
            for (int q = 0; q < kStage; ++q) {
                if (q0 + q * kThreads >= n) break;
                const int p = rswz(pos[q]);
                sv[0][p] = (Td[q] > 1e-10 && val[q] >= 0.0f) ? val[q] : -1.0f;  // emc_atomic.cu:190, :56
                sv[1][p] = (float)Td[q];
            }
        }
        __syncthreads();
        // 2. positions [k0, k1); lv = the base voxel of the current run
        int lv = 0;
        if (k0 < n) {  // the voxel whose run holds k0: the last lv with svb[lv] <= k0
            int lo = 0, hi = kBrick3 - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (svb[mid] <= k0) lo = mid;
                else hi = mid - 1;
            }
            lv = lo;
        }
        float acc[2 * NC], own[2 * NC], mine[2 * NC];
#pragma unroll
        for (int c = 0; c < 2 * NC; ++c) acc[c] = own[c] = mine[c] = 0.0f;
        int own_lv = -1, my_tail = -1;
        auto close_run = [&](int kend) {  // the run of lv ends at kend (exclusive)
            const bool open_left = svb[lv] < k0, open_right = svb[lv + 1] > kend;
            if (open_right) {  // continues into the next thread: handed over after step 2
#pragma unroll
                for (int c = 0; c < 2 * NC; ++c) mine[c] = acc[c];
                my_tail = lv;
            } else if (open_left) {  // this thread holds the run's last position: adds the pieces later
#pragma unroll
                for (int c = 0; c < 2 * NC; ++c) own[c] = acc[c];
                own_lv = lv;
            } else {  // the whole run: only this thread touches lv in this unit
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    cM[c][lv] += acc[c];
                    cW[c][lv] += acc[NC + c];
                }
            }
#pragma unroll
            for (int c = 0; c < 2 * NC; ++c) acc[c] = 0.0f;
        };
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int k = k0 + i;
            if (k >= k1) break;
            if (k >= svb[lv + 1]) {
                close_run(k);
                while (svb[lv + 1] <= k) ++lv;  // skip base voxels without pairs
            }
            const int p = rswz(k);
            const float v = sv[0][p];
            if (v < 0.0f) continue;
            const float Ts = sv[1][p];
            if constexpr (INTERP == EMC_INTERP_NEAREST) {
                acc[0] += v;
                acc[1] += Ts;
            } else {
                const float gw[2][3] = {{1.0f - fr[i][0], 1.0f - fr[i][1], 1.0f - fr[i][2]},
                                        {fr[i][0], fr[i][1], fr[i][2]}};
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const float w = gw[c & 1][0] * gw[(c >> 1) & 1][1] * gw[c >> 2][2];
                    acc[c] += w * v;
                    acc[NC + c] += w * Ts;
                }
            }
        }
        if (k0 < k1) close_run(k1);
        __syncthreads();  // the staging is consumed: the tails take its space
        tail_lv[tid] = my_tail;
        if (my_tail >= 0)
#pragma unroll
            for (int c = 0; c < 2 * NC; ++c) tail[c][tid] = mine[c];
        __syncthreads();
        if (own_lv >= 0) {  // the pieces of earlier threads, in thread order, then this thread's
            int t0 = tid;
            while (t0 > 0 && tail_lv[t0 - 1] == own_lv) --t0;
            float sum[2 * NC];
#pragma unroll
            for (int c = 0; c < 2 * NC; ++c) sum[c] = 0.0f;
            for (int t = t0; t < tid; ++t)
#pragma unroll
                for (int c = 0; c < 2 * NC; ++c) sum[c] += tail[c][t];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                cM[c][own_lv] += sum[c] + own[c];
                cW[c][own_lv] += sum[NC + c] + own[NC + c];
            }
        }
    }
    __syncthreads();
    // 3. each halo voxel gathers the corner sums of the base voxels below it, in corner order
    float *pM = part + (size_t)b * 2 * kHalo3, *pW = pM + kHalo3;
    for (int hv = tid; hv < kHalo3; hv += kThreads) {
        const int hx = hv % kHalo, hy = (hv / kHalo) % kHalo, hz = hv / (kHalo * kHalo);
        float sm = 0.0f, sw = 0.0f;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int bx = hx - (c & 1), by = hy - ((c >> 1) & 1), bz = hz - (c >> 2);
            if (bx < 0 || by < 0 || bz < 0 || bx >= kBrick || by >= kBrick || bz >= kBrick) continue;
            const int lb = (bz * kBrick + by) * kBrick + bx;
            sm += cM[c][lb];
            sw += cW[c][lb];
        }
        pM[hv] = sm;
        pW[hv] = sw;
    }
}

// Step 4a: group g sums the partials of its blocks [gu0, gu1) (consecutive blocks of one brick), in
// block order.  groups: [g][4] = key, first block, end block, 0 (blocks relative to the chunk).
__global__ __launch_bounds__(kThreads) void k_bin_group(const int32_t *__restrict__ groups,
                                                        const float *__restrict__ part, float *__restrict__ gpart)
{
    const int gi = blockIdx.x;
    const int u0 = groups[4 * gi + 1], u1 = groups[4 * gi + 2];
    float *out = gpart + (size_t)gi * 2 * kHalo3;
    for (int e = threadIdx.x; e < 2 * kHalo3; e += kThreads) {
        float acc = 0.0f;
        int u = u0;
        for (; u + 4 <= u1; u += 4) {  // four independent loads, added in order
            const float a = part[(size_t)u * 2 * kHalo3 + e], b = part[(size_t)(u + 1) * 2 * kHalo3 + e];
            const float c = part[(size_t)(u + 2) * 2 * kHalo3 + e], d = part[(size_t)(u + 3) * 2 * kHalo3 + e];
            acc += a;
            acc += b;
            acc += c;
            acc += d;
        }
        for (; u < u1; ++u) acc += part[(size_t)u * 2 * kHalo3 + e];
        out[e] = acc;
    }
}

// Step 4b: every voxel adds the groups of the bricks whose halo covers it: its own brick and, on
// a brick's low face, the neighbour below on that axis (whose halo index there is kBrick); in a
// fixed order.  gbeg: [key - key0] first group of each brick of the chunk (kBricks + 1 entries).
__global__ __launch_bounds__(kThreads) void k_insert_reduce(const float *__restrict__ gpart,
                                                            const int32_t *__restrict__ gbeg, BinGeom g,
                                                            float *__restrict__ Mnum, float *__restrict__ Wt,
                                                            bool first)
{
    const int64_t v = blockIdx.x * (int64_t)kThreads + threadIdx.x;
    const int L = g.L;
    if (v >= (int64_t)L * L * L) return;
    const int x = (int)(v % L), y = (int)((v / L) % L), z = (int)(v / ((int64_t)L * L));
    const int bx = x / kBrick, by = y / kBrick, bz = z / kBrick;
    const int lx = x % kBrick, ly = y % kBrick, lz = z % kBrick;
    const int nx = (lx == 0 && bx > 0) ? 2 : 1, ny = (ly == 0 && by > 0) ? 2 : 1, nz = (lz == 0 && bz > 0) ? 2 : 1;
    float sm = 0.0f, sw = 0.0f;
    for (int dz = 0; dz < nz; ++dz)
        for (int dy = 0; dy < ny; ++dy)
            for (int dx = 0; dx < nx; ++dx) {
                const int key = ((bz - dz) * g.nb + (by - dy)) * g.nb + (bx - dx);
                const int hv = ((dz ? kBrick : lz) * kHalo + (dy ? kBrick : ly)) * kHalo + (dx ? kBrick : lx);
                for (int gi = gbeg[key]; gi < gbeg[key + 1]; ++gi) {
                    sm += gpart[(size_t)gi * 2 * kHalo3 + hv];
                    sw += gpart[(size_t)gi * 2 * kHalo3 + kHalo3 + hv];
                }
            }
    if (first) {  // the pass's first rotation chunk: the sums start here (no separate zero fill)
        Mnum[v] = sm;
        Wt[v] = sw;
    } else {
        Mnum[v] += sm;
        Wt[v] += sw;
    }
}

BinGeom geom_of(const emc_ctx *c)
{
    BinGeom g;
    g.L = c->L;
    g.nb = (c->L + kBrick - 1) / kBrick;
    g.S = c->S;
    g.Pa = c->Pa;
    g.Pp = c->Pp;
    g.chunk = c->chunk;
    g.interp = c->opt.interp;
    int jb = 1;
    while ((1 << jb) < c->Pa) ++jb;
    g.jbits = jb;
    return g;
}

}  // namespace

// Builds the pair lists, units, groups and their device tables.  Allocates (setup time only).
hipError_t build_bins(emc_ctx *c, std::string &err)
{
    hipStream_t st = c->stream;
    const BinGeom g = geom_of(c);
    const int64_t n = (int64_t)c->S * c->Pa;
    if (n >= ((int64_t)1 << 31) || ((int64_t)c->S << g.jbits) > ((int64_t)1 << 32)) {
        err = "binned insertion supports S * active pixels < 2^31 and S * 2^ceil(log2 P_act) <= 2^32 "
              "(use insert = ATOMIC)";
        return hipErrorInvalidValue;
    }
    const int nch = (c->S + c->chunk - 1) / c->chunk;
    const uint32_t nbr = (uint32_t)g.nb * g.nb * g.nb;
    const uint32_t nkeys = (uint32_t)nch * nbr;
    int end_bit = 1;
    while (end_bit < 32 && ((uint64_t)1 << end_bit) <= nkeys) ++end_bit;
    uint32_t *kin = nullptr, *vin = nullptr, *kout = nullptr, *pout = nullptr;
    void *tmp = nullptr;
    int32_t *cnt = nullptr;
    size_t tmp_bytes = 0;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    auto release = [](auto *&p) {
        if (p) (void)hipFree(p);
        p = nullptr;
    };
    release(c->bin_vals);
    release(c->bin_perm);
    release(c->bin_vbeg);
    release(c->bin_frac);
    release(c->bin_units);
    release(c->bin_blocks);
    release(c->bin_groups);
    release(c->bin_gbeg);
    release(c->bin_part);
    release(c->bin_gpart);
    chk(hipMalloc((void **)&kin, sizeof(uint32_t) * n));
    chk(hipMalloc((void **)&vin, sizeof(uint32_t) * n));
    chk(hipMalloc((void **)&kout, sizeof(uint32_t) * n));
    chk(hipMalloc((void **)&c->bin_vals, sizeof(uint32_t) * n));
    chk(hipMalloc((void **)&cnt, sizeof(int32_t) * 2 * (size_t)nkeys));
    const int grid1 = (int)std::min<int64_t>((n + 255) / 256, 65536);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_bin_keys, dim3(grid1), dim3(256), 0, st, c->rotm, c->axyz, g, n, nkeys, kin, vin);
        chk(hipGetLastError());
    }
    if (e == hipSuccess)
        chk(rocprim::radix_sort_pairs(nullptr, tmp_bytes, kin, kout, vin, c->bin_vals, (size_t)n, 0, end_bit, st));
    if (e == hipSuccess) chk(hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 16)));
    if (e == hipSuccess)
        chk(rocprim::radix_sort_pairs(tmp, tmp_bytes, kin, kout, vin, c->bin_vals, (size_t)n, 0, end_bit, st));
    if (e == hipSuccess) chk(hipMemsetAsync(cnt, 0, sizeof(int32_t) * 2 * (size_t)nkeys, st));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_bin_count, dim3(grid1), dim3(256), 0, st, kout, n, nkeys, cnt);
        chk(hipGetLastError());
    }
    std::vector<int32_t> h(2 * (size_t)nkeys, 0);
    if (e == hipSuccess) chk(hipMemcpyAsync(h.data(), cnt, sizeof(int32_t) * h.size(), hipMemcpyDeviceToHost, st));
    if (e == hipSuccess) chk(hipStreamSynchronize(st));
    // units (key, begin, end, 0): at most kUnit pairs, key order (chunk-major); blocks (key, first
    // unit, end unit, 0): at most kUnitsPerBlock consecutive units of one brick, one insertion
    // workgroup each, numbered per chunk (unit indices global); groups (key, first block, end
    // block, 0): at most kGroup consecutive blocks of one brick, numbered per chunk, with each
    // brick's first group in gbeg
    std::vector<int32_t> units, blocks, groups, gbeg;
    c->bin_chunk_blocks.assign(nch + 1, 0);
    c->bin_chunk_groups.assign(nch + 1, 0);
    int max_blocks = 0, max_groups = 0;
    for (int ch = 0; ch < nch && e == hipSuccess; ++ch) {
        const int b_first = (int)(blocks.size() / 4), g_first = (int)(groups.size() / 4);
        for (uint32_t br = 0; br < nbr; ++br) {
            const uint32_t k = (uint32_t)ch * nbr + br;
            gbeg.push_back((int32_t)(groups.size() / 4) - g_first);
            const int32_t kb = h[2 * k], ke = h[2 * k + 1];
            const int ub = (int)(units.size() / 4);
            for (int32_t t = kb; t < ke; t += kUnit) {
                units.push_back((int32_t)k);
                units.push_back(t);
                units.push_back(std::min(ke, t + kUnit));
                units.push_back(0);
            }
            const int ue = (int)(units.size() / 4);
            const int bb = (int)(blocks.size() / 4) - b_first;
            for (int u = ub; u < ue; u += kUnitsPerBlock) {
                blocks.push_back((int32_t)k);
                blocks.push_back(u);
                blocks.push_back(std::min(ue, u + kUnitsPerBlock));
                blocks.push_back(0);
            }
            const int be = (int)(blocks.size() / 4) - b_first;
            for (int bl = bb; bl < be; bl += kGroup) {
                groups.push_back((int32_t)k);
                groups.push_back(bl);
                groups.push_back(std::min(be, bl + kGroup));
                groups.push_back(0);
            }
        }
        gbeg.push_back((int32_t)(groups.size() / 4) - g_first);
        c->bin_chunk_blocks[ch + 1] = (int)(blocks.size() / 4);
        c->bin_chunk_groups[ch + 1] = (int)(groups.size() / 4);
        max_blocks = std::max(max_blocks, c->bin_chunk_blocks[ch + 1] - c->bin_chunk_blocks[ch]);
        max_groups = std::max(max_groups, c->bin_chunk_groups[ch + 1] - c->bin_chunk_groups[ch]);
    }
    const uint32_t n_units = (uint32_t)(units.size() / 4);
    if (e == hipSuccess && n_units > 0) {
        chk(hipMalloc((void **)&c->bin_units, sizeof(int32_t) * units.size()));
        chk(hipMalloc((void **)&c->bin_blocks, sizeof(int32_t) * blocks.size()));
        chk(hipMalloc((void **)&c->bin_groups, sizeof(int32_t) * groups.size()));
        chk(hipMalloc((void **)&c->bin_gbeg, sizeof(int32_t) * gbeg.size()));
        chk(hipMalloc((void **)&c->bin_part, sizeof(float) * 2 * kHalo3 * (size_t)max_blocks));
        chk(hipMalloc((void **)&c->bin_gpart, sizeof(float) * 2 * kHalo3 * (size_t)max_groups));
        chk(hipMalloc((void **)&c->bin_perm, sizeof(uint16_t) * n));
        chk(hipMalloc((void **)&c->bin_vbeg, sizeof(uint16_t) * kVbeg * (size_t)n_units));
        if (c->opt.interp == EMC_INTERP_TRILINEAR)
            chk(hipMalloc((void **)&c->bin_frac, sizeof(float) * 3 * kUnit * (size_t)n_units));
        if (e == hipSuccess)
            chk(hipMemcpy(c->bin_units, units.data(), sizeof(int32_t) * units.size(), hipMemcpyHostToDevice));
        if (e == hipSuccess)
            chk(hipMemcpy(c->bin_blocks, blocks.data(), sizeof(int32_t) * blocks.size(), hipMemcpyHostToDevice));
        if (e == hipSuccess)
            chk(hipMemcpy(c->bin_groups, groups.data(), sizeof(int32_t) * groups.size(), hipMemcpyHostToDevice));
        if (e == hipSuccess) chk(hipMemcpy(c->bin_gbeg, gbeg.data(), sizeof(int32_t) * gbeg.size(), hipMemcpyHostToDevice));
        // within each unit, rank the pairs by base voxel (stable: then by (s, j))
        const uint64_t nkeys2 = (uint64_t)n_units * kBrick3;
        int end_bit2 = 1;
        while (end_bit2 < 32 && ((uint64_t)1 << end_bit2) <= nkeys2) ++end_bit2;
        if (nkeys2 >= ((uint64_t)1 << 32)) chk(hipErrorInvalidValue);
        // pairs outside every unit (no valid voxel) keep the largest key and stay at the end
        const uint32_t sentinel = (uint32_t)(((uint64_t)1 << end_bit2) - 1);
        if (e == hipSuccess) chk(hipMemsetD32Async((hipDeviceptr_t)kin, (int)sentinel, (size_t)n, st));
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_unit_keys, dim3(n_units), dim3(kThreads), 0, st, c->bin_units, c->bin_vals, c->rotm,
                               c->axyz, g, kin);
            chk(hipGetLastError());
        }
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_iota, dim3(grid1), dim3(256), 0, st, vin, n);
            chk(hipGetLastError());
        }
        chk(hipMalloc((void **)&pout, sizeof(uint32_t) * n));
        size_t tmp2 = 0;
        if (e == hipSuccess)
            chk(rocprim::radix_sort_pairs(nullptr, tmp2, kin, kout, vin, pout, (size_t)n, 0, end_bit2, st));
        if (e == hipSuccess && tmp2 > tmp_bytes) {
            (void)hipFree(tmp);
            tmp = nullptr;
            chk(hipMalloc(&tmp, tmp2));
            tmp_bytes = tmp2;
        }
        if (e == hipSuccess)
            chk(rocprim::radix_sort_pairs(tmp, tmp_bytes, kin, kout, vin, pout, (size_t)n, 0, end_bit2, st));
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_unit_perm, dim3(n_units), dim3(kThreads), 0, st, c->bin_units, kout, pout, c->bin_vals,
                               c->rotm, c->axyz, g, c->bin_perm, c->bin_vbeg, c->bin_frac);
            chk(hipGetLastError());
        }
        if (e == hipSuccess) chk(hipStreamSynchronize(st));
    }
    (void)hipFree(kin);
    (void)hipFree(vin);
    (void)hipFree(kout);
    (void)hipFree(pout);
    (void)hipFree(tmp);
    (void)hipFree(cnt);
    if (e != hipSuccess && err.empty()) err = std::string("build_bins: ") + hipGetErrorString(e);
    c->bins_valid = e == hipSuccess;
    return e;
}

hipError_t launch_insert_binned(const emc_ctx *c, int s0, int cs, hipStream_t st)
{
    (void)cs;
    const int k = s0 / c->chunk;
    const int b0 = c->bin_chunk_blocks[k], b1 = c->bin_chunk_blocks[k + 1];
    const int g0 = c->bin_chunk_groups[k], g1 = c->bin_chunk_groups[k + 1];
    if (b1 <= b0) {  // no pairs in this chunk; the first chunk still starts the sums at zero
        hipError_t e = hipSuccess;
        if (k == 0) e = launch_fill_f32(c->Mnum, 0.0f, c->N3, st);
        if (e == hipSuccess && k == 0) e = launch_fill_f32(c->Wt, 0.0f, c->N3, st);
        return e;
    }
    const BinGeom g = geom_of(c);
    const int32_t *blocks = c->bin_blocks + 4 * (size_t)b0;
    if (c->opt.interp == EMC_INTERP_NEAREST)
        hipLaunchKernelGGL(k_insert_binned<EMC_INTERP_NEAREST>, dim3(b1 - b0), dim3(kThreads), 0, st, blocks,
                           c->bin_units, c->bin_vals, c->bin_perm, c->bin_vbeg, nullptr, c->U, s0, c->T, g, c->bin_part);
    else
        hipLaunchKernelGGL(k_insert_binned<EMC_INTERP_TRILINEAR>, dim3(b1 - b0), dim3(kThreads), 0, st, blocks,
                           c->bin_units, c->bin_vals, c->bin_perm, c->bin_vbeg, c->bin_frac, c->U, s0, c->T, g,
                           c->bin_part);
    hipLaunchKernelGGL(k_bin_group, dim3(g1 - g0), dim3(kThreads), 0, st, c->bin_groups + 4 * (size_t)g0, c->bin_part,
                       c->bin_gpart);
    const int nbr = g.nb * g.nb * g.nb;
    hipLaunchKernelGGL(k_insert_reduce, dim3((unsigned)((c->N3 + kThreads - 1) / kThreads)), dim3(kThreads), 0, st,
                       c->bin_gpart, c->bin_gbeg + (size_t)k * (nbr + 1), g, c->Mnum, c->Wt, k == 0);
    return hipGetLastError();
}

}  // namespace emc
