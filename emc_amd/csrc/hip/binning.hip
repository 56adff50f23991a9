// Maximize, scatter half, binned: insertion without global-atomic contention.
//
// The geometry of insertion (which voxels a rotated pixel touches, with which trilinear weights)
// depends only on the rotation list, the detector coordinates and the mask, not on the
// iteration.  So once per geometry the valid (rotation s, active pixel j) pairs are sorted by the
// model brick (kBrick^3 voxels) that holds their base voxel, per rotation chunk, with a stable
// radix sort (rocPRIM).  Every iteration then runs one workgroup per slice of a brick's list:
// it accumulates U[s, j] and T_s with LDS atomics into a (kBrick+1)^3 private copy of the brick
// (the +1 is the trilinear halo) and flushes it to the model numerator and weight once.  The
// workgroup first stages its items' U[s, j] and T_s in LDS in (s, j) order, where the reads of U
// are runs of consecutive pixels; a second, per-unit permutation (by base voxel) then drives the
// accumulation, so that runs of one voxel are summed in registers before the LDS atomics.  This
// replaces the 16 global atomics per (s, j) of the reference's insert_slices_kernel
// (emc_atomic.cu:33-85, 183-200).
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
constexpr int kHalo = kBrick + 1;
constexpr int kUnit = 4096;  // items per workgroup
constexpr int kStage = 8;    // items per thread whose loads the insertion issues together

struct BinGeom {
    int L, nb, S, Pa, Pp, chunk, interp;
};

// Voxel cell of active pixel j under rotation matrix m (nearest: the voxel; trilinear: the base
// corner and fractions); false when out of bounds.
__device__ __forceinline__ bool item_cell_m(const float *m, const float *__restrict__ axyz, const BinGeom &g, int j,
                                            int &bx, int &by, int &bz, float &fx, float &fy, float &fz)
{
    float ux, uy, uz;
    rot_point(m, axyz[j], axyz[g.Pp + j], axyz[2 * g.Pp + j], model_c(g.L), ux, uy, uz);
    if (g.interp == EMC_INTERP_NEAREST) {
        bx = nearest(ux);
        by = nearest(uy);
        bz = nearest(uz);
        fx = fy = fz = 0.0f;
        return bx >= 0 && bx < g.L && by >= 0 && by < g.L && bz >= 0 && bz < g.L;  // emc_atomic.cu:67-69
    }
    return tri_axis(ux, g.L, bx, fx) && tri_axis(uy, g.L, by, fy) && tri_axis(uz, g.L, bz, fz);
}

__device__ __forceinline__ bool item_cell(const float *__restrict__ rotm, const float *__restrict__ axyz,
                                          const BinGeom &g, int s, int j, int &bx, int &by, int &bz, float &fx,
                                          float &fy, float &fz)
{
    float m[9];
    load_rot(rotm, s, m);
    return item_cell_m(m, axyz, g, j, bx, by, bz, fx, fy, fz);
}

__global__ void k_bin_keys(const float *__restrict__ rotm, const float *__restrict__ axyz, BinGeom g,
                           int64_t n, uint32_t invalid_key, uint32_t *__restrict__ keys, uint32_t *__restrict__ vals)
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
        vals[idx] = (uint32_t)idx;
    }
}

__global__ void k_bin_count(const uint32_t *__restrict__ keys, int64_t n, uint32_t nkeys, int32_t *__restrict__ counts)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t k = keys[t];
        if (k < nkeys && (t == 0 || keys[t - 1] != k)) counts[2 * k] = (int32_t)t;       // first position
        if (k < nkeys && (t == n - 1 || keys[t + 1] != k)) counts[2 * k + 1] = (int32_t)(t + 1);  // end
    }
}

__global__ void k_iota(uint32_t *__restrict__ v, int64_t n)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        v[t] = (uint32_t)t;
}

// Second sort key of the items of unit u: u * 512 + the base voxel's index inside its 8^3 brick.
__global__ __launch_bounds__(256) void k_unit_keys(const int32_t *__restrict__ units, const uint32_t *__restrict__ vals,
                                                   const float *__restrict__ rotm, const float *__restrict__ axyz,
                                                   BinGeom g, uint32_t *__restrict__ keys)
{
    const int u = blockIdx.x;
    const int begin = units[4 * u + 1], end = units[4 * u + 2];
    for (int t = begin + threadIdx.x; t < end; t += 256) {
        const uint32_t idx = vals[t];
        const int s = (int)(idx / (uint32_t)g.Pa), j = (int)(idx % (uint32_t)g.Pa);
        int bx = 0, by = 0, bz = 0;
        float fx, fy, fz;
        item_cell(rotm, axyz, g, s, j, bx, by, bz, fx, fy, fz);  // valid: only valid pairs are listed
        const int lv = ((bz % kBrick) * kBrick + by % kBrick) * kBrick + bx % kBrick;
        keys[t] = (uint32_t)u * (kBrick * kBrick * kBrick) + (uint32_t)lv;
    }
}

// The insertion's per-item stream of unit u, in base-voxel order: item q of the unit belongs to
// thread q / per (per = ceil(n / 256) items each, a contiguous run) and is stored at
// soff + (q % per) * 256 + q / per, so that at each step the 256 threads read 256 consecutive
// entries.  Each entry holds the item's position in vals (ipos), its base voxel inside the unit's
// halo brick (vox; 0xFFFF marks the padding after the last item) and, for trilinear, its
// fractions (frac).  Iteration-independent, computed with the bins.
__global__ __launch_bounds__(256) void k_item_geom(const int32_t *__restrict__ units, const uint32_t *__restrict__ vals,
                                                   const uint32_t *__restrict__ perm, const float *__restrict__ rotm,
                                                   const float *__restrict__ axyz, BinGeom g,
                                                   uint32_t *__restrict__ ipos, uint16_t *__restrict__ vox,
                                                   float *__restrict__ frac)
{
    const int u = blockIdx.x;
    const int brick = units[4 * u] % (g.nb * g.nb * g.nb);
    const int begin = units[4 * u + 1], end = units[4 * u + 2];
    const int64_t soff = units[4 * u + 3];
    const int n = end - begin, per = (n + 255) / 256;
    const int ox = (brick % g.nb) * kBrick, oy = ((brick / g.nb) % g.nb) * kBrick, oz = (brick / (g.nb * g.nb)) * kBrick;
    for (int q = threadIdx.x; q < per * 256; q += 256) {
        const int64_t st = soff + (int64_t)(q % per) * 256 + q / per;
        if (q >= n) {
            ipos[st] = (uint32_t)begin;
            vox[st] = 0xFFFF;
            continue;
        }
        const uint32_t pos = perm[begin + q];
        const uint32_t idx = vals[pos];
        const int s = (int)(idx / (uint32_t)g.Pa), j = (int)(idx % (uint32_t)g.Pa);
        int bx = 0, by = 0, bz = 0;
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        item_cell(rotm, axyz, g, s, j, bx, by, bz, fx, fy, fz);  // valid: only valid pairs are listed
        ipos[st] = pos;
        vox[st] = (uint16_t)(((bz - oz) * kHalo + (by - oy)) * kHalo + (bx - ox));
        if (frac) {
            frac[3 * st] = fx;
            frac[3 * st + 1] = fy;
            frac[3 * st + 2] = fz;
        }
    }
}

template <int INTERP>
__global__ __launch_bounds__(256) void k_insert_binned(const int32_t *__restrict__ units, const uint32_t *__restrict__ vals,
                                                       const uint32_t *__restrict__ perm, const uint16_t *__restrict__ vox,
                                                       const float *__restrict__ frac, const float *__restrict__ U,
                                                       int s0, const double *__restrict__ T, BinGeom g,
                                                       float *__restrict__ Mnum, float *__restrict__ Wt)
{
    __shared__ float accM[kHalo * kHalo * kHalo], accW[kHalo * kHalo * kHalo];
    __shared__ float sval[kUnit], sts[kUnit];
    const int u = blockIdx.x;
    const int brick = units[4 * u] % (g.nb * g.nb * g.nb);
    const int begin = units[4 * u + 1], end = units[4 * u + 2];
    const int ox = (brick % g.nb) * kBrick, oy = ((brick / g.nb) % g.nb) * kBrick, oz = (brick / (g.nb * g.nb)) * kBrick;
    for (int k = threadIdx.x; k < kHalo * kHalo * kHalo; k += 256) {
        accM[k] = 0.0f;
        accW[k] = 0.0f;
    }
    // Stage U[s, j] and T_s of the unit's items in (s, j) order, where the reads of U are runs
    // of consecutive pixels; -1 marks an item the insertion skips.  kStage items per thread and
    // round, their loads issued together.
    const int n = end - begin;
    for (int k0 = threadIdx.x; k0 < n; k0 += 256 * kStage) {
        uint32_t idx[kStage];
#pragma unroll
        for (int q = 0; q < kStage; ++q) idx[q] = vals[begin + min(k0 + q * 256, n - 1)];  // past the end: a valid item, not stored
        double Td[kStage];
        float val[kStage];
#pragma unroll
        for (int q = 0; q < kStage; ++q) {
            const int s = (int)(idx[q] / (uint32_t)g.Pa), j = (int)(idx[q] % (uint32_t)g.Pa);
            Td[q] = T[s];
            val[q] = U[(size_t)(s - s0) * g.Pp + j];
        }
#pragma unroll
        for (int q = 0; q < kStage; ++q) {
            if (k0 + q * 256 >= n) break;
            // emc_atomic.cu:190 (T_s > 1e-10) and emc_atomic.cu:56 (U >= 0)
            sval[k0 + q * 256] = (Td[q] > 1e-10 && val[q] >= 0.0f) ? val[q] : -1.0f;
            sts[k0 + q * 256] = (float)Td[q];
        }
    }
    __syncthreads();
    // perm lists the unit's items by base voxel (with the voxel and the trilinear fractions
    // beside it), so each thread walks a contiguous run of them and sums the corner contributions
    // in registers while the base voxel repeats; only at a change of voxel (and at the end) do
    // the sums go to the LDS brick with atomics.
    constexpr int NC = INTERP == EMC_INTERP_NEAREST ? 1 : 8;
    const int per = (end - begin + 255) / 256;
    const int64_t soff = units[4 * u + 3];
    float sm[NC], sw[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) sm[k] = sw[k] = 0.0f;
    int cur = -1;
    auto flush = [&]() {
        if (cur < 0) return;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            const int l = cur + ((k >> 2) * kHalo + ((k >> 1) & 1)) * kHalo + (k & 1);
            atomicAdd(accM + l, sm[k]);
            atomicAdd(accW + l, sw[k]);
            sm[k] = sw[k] = 0.0f;
        }
    };
    // kStage steps of the stream are loaded together, then summed in order
    for (int k0 = 0; k0 < per; k0 += kStage) {
        int lv[kStage];
        uint32_t pv[kStage];
        float fv[kStage][3];
#pragma unroll
        for (int q = 0; q < kStage; ++q) {
            const int64_t t = soff + (int64_t)(k0 + q) * 256 + threadIdx.x;
            const bool in = k0 + q < per;
            const int64_t tc = in ? t : soff + threadIdx.x;  // past the stream's end: entry 0, not used
            lv[q] = in ? (int)vox[tc] : 0xFFFF;
            pv[q] = perm[tc];
#pragma unroll
            for (int a = 0; a < 3; ++a) fv[q][a] = INTERP == EMC_INTERP_TRILINEAR ? frac[3 * tc + a] : 0.0f;
        }
        bool done = false;
#pragma unroll
        for (int q = 0; q < kStage; ++q) {
            const int l = lv[q];
            if (l == 0xFFFF) {  // padding: this thread's run has ended
                done = true;
                break;
            }
            const float val = sval[pv[q] - begin];
            if (val < 0.0f) continue;
            const float Ts = sts[pv[q] - begin];
            if (l != cur) {
                flush();
                cur = l;
            }
            if (INTERP == EMC_INTERP_NEAREST) {
                sm[0] += val;
                sw[0] += Ts;
            } else {
                const float fx = fv[q][0], fy = fv[q][1], fz = fv[q][2];
                const float gw[2][3] = {{1.0f - fx, 1.0f - fy, 1.0f - fz}, {fx, fy, fz}};
#pragma unroll
                for (int k = 0; k < NC; ++k) {
                    const float w = gw[k & 1][0] * gw[(k >> 1) & 1][1] * gw[k >> 2][2];
                    sm[k] += w * val;
                    sw[k] += w * Ts;
                }
            }
        }
        if (done) break;
    }
    flush();
    __syncthreads();
    for (int k = threadIdx.x; k < kHalo * kHalo * kHalo; k += 256) {
        const float a = accM[k], b = accW[k];
        if (a == 0.0f && b == 0.0f) continue;
        const int x = ox + k % kHalo, y = oy + (k / kHalo) % kHalo, z = oz + k / (kHalo * kHalo);
        if (x >= g.L || y >= g.L || z >= g.L) continue;
        const size_t v = ((size_t)z * g.L + y) * g.L + x;
        atomicAdd(Mnum + v, a);
        atomicAdd(Wt + v, b);
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
    return g;
}

}  // namespace

// Builds c->bin_vals / c->bin_units / c->bin_chunk_units.  Allocates (setup time only).
hipError_t build_bins(emc_ctx *c, std::string &err)
{
    hipStream_t st = c->stream;
    const BinGeom g = geom_of(c);
    const int64_t n = (int64_t)c->S * c->Pa;
    if (n >= ((int64_t)1 << 31)) {
        err = "binned insertion supports S * active pixels < 2^31";
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
    if (c->bin_vals) (void)hipFree(c->bin_vals);
    if (c->bin_perm) (void)hipFree(c->bin_perm);
    if (c->bin_vox) (void)hipFree(c->bin_vox);
    if (c->bin_frac) (void)hipFree(c->bin_frac);
    if (c->bin_units) (void)hipFree(c->bin_units);
    c->bin_vals = nullptr;
    c->bin_perm = nullptr;
    c->bin_vox = nullptr;
    c->bin_frac = nullptr;
    c->bin_units = nullptr;
    chk(hipMalloc((void **)&kin, sizeof(uint32_t) * n));
    chk(hipMalloc((void **)&vin, sizeof(uint32_t) * n));
    chk(hipMalloc((void **)&kout, sizeof(uint32_t) * n));
    chk(hipMalloc((void **)&c->bin_vals, sizeof(uint32_t) * n));
    chk(hipMalloc((void **)&cnt, sizeof(int32_t) * 2 * (size_t)nkeys));
    if (e == hipSuccess) {
        const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
        hipLaunchKernelGGL(k_bin_keys, dim3(blocks), dim3(256), 0, st, c->rotm, c->axyz, g, n, nkeys, kin, vin);
        chk(hipGetLastError());
    }
    if (e == hipSuccess)
        chk(rocprim::radix_sort_pairs(nullptr, tmp_bytes, kin, kout, vin, c->bin_vals, (size_t)n, 0, end_bit, st));
    if (e == hipSuccess) chk(hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 16)));
    if (e == hipSuccess)
        chk(rocprim::radix_sort_pairs(tmp, tmp_bytes, kin, kout, vin, c->bin_vals, (size_t)n, 0, end_bit, st));
    if (e == hipSuccess) chk(hipMemsetAsync(cnt, 0, sizeof(int32_t) * 2 * (size_t)nkeys, st));
    if (e == hipSuccess) {
        const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
        hipLaunchKernelGGL(k_bin_count, dim3(blocks), dim3(256), 0, st, kout, n, nkeys, cnt);
        chk(hipGetLastError());
    }
    std::vector<int32_t> h(2 * (size_t)nkeys, 0);
    if (e == hipSuccess) chk(hipMemcpyAsync(h.data(), cnt, sizeof(int32_t) * h.size(), hipMemcpyDeviceToHost, st));
    if (e == hipSuccess) chk(hipStreamSynchronize(st));
    // work units (key, begin, end, offset of the insertion stream), at most kUnit items each,
    // ordered by key = chunk-major; a unit's stream holds ceil(items / 256) * 256 entries
    std::vector<int32_t> units;
    int64_t ns_total = 0;
    c->bin_chunk_units.assign(nch + 1, 0);
    for (uint32_t k = 0; k < nkeys; ++k) {
        const int ch = (int)(k / nbr);
        const int32_t b = h[2 * k], en = h[2 * k + 1];
        for (int32_t t = b; t < en; t += kUnit) {
            const int32_t t1 = std::min(en, t + kUnit);
            units.push_back((int32_t)k);
            units.push_back(t);
            units.push_back(t1);
            units.push_back((int32_t)ns_total);
            ns_total += (int64_t)(t1 - t + 255) / 256 * 256;
        }
        c->bin_chunk_units[ch + 1] = (int)(units.size() / 4);
    }
    if (ns_total >= ((int64_t)1 << 31)) chk(hipErrorInvalidValue);
    for (int ch = 1; ch <= nch; ++ch) c->bin_chunk_units[ch] = std::max(c->bin_chunk_units[ch], c->bin_chunk_units[ch - 1]);
    if (e == hipSuccess && !units.empty()) {
        chk(hipMalloc((void **)&c->bin_units, sizeof(int32_t) * units.size()));
        if (e == hipSuccess)
            chk(hipMemcpy(c->bin_units, units.data(), sizeof(int32_t) * units.size(), hipMemcpyHostToDevice));
        // within each unit, list the items' positions by base voxel (stable: then by (s, j)), so
        // the insertion can sum runs of one voxel in registers; bin_vals keeps the (s, j) order
        const uint32_t n_units = (uint32_t)(units.size() / 4);
        const uint64_t nkeys2 = (uint64_t)n_units * kBrick * kBrick * kBrick;
        int end_bit2 = 1;
        while (end_bit2 < 32 && ((uint64_t)1 << end_bit2) <= nkeys2) ++end_bit2;
        if (nkeys2 >= ((uint64_t)1 << 32)) chk(hipErrorInvalidValue);
        // items outside every unit (no valid voxel) keep the largest key and stay at the end
        const uint32_t sentinel = (uint32_t)(((uint64_t)1 << end_bit2) - 1);
        if (e == hipSuccess) chk(hipMemsetD32Async((hipDeviceptr_t)kin, (int)sentinel, (size_t)n, st));
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_unit_keys, dim3(n_units), dim3(256), 0, st, c->bin_units, c->bin_vals, c->rotm, c->axyz,
                               g, kin);
            chk(hipGetLastError());
        }
        if (e == hipSuccess) {
            const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
            hipLaunchKernelGGL(k_iota, dim3(blocks), dim3(256), 0, st, vin, n);
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
        chk(hipMalloc((void **)&c->bin_perm, sizeof(uint32_t) * ns_total));
        chk(hipMalloc((void **)&c->bin_vox, sizeof(uint16_t) * ns_total));
        if (c->opt.interp == EMC_INTERP_TRILINEAR) chk(hipMalloc((void **)&c->bin_frac, sizeof(float) * 3 * ns_total));
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_item_geom, dim3(n_units), dim3(256), 0, st, c->bin_units, c->bin_vals, pout, c->rotm,
                               c->axyz, g, c->bin_perm, c->bin_vox, c->bin_frac);
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
    const int u0 = c->bin_chunk_units[k], u1 = c->bin_chunk_units[k + 1];
    if (u1 <= u0) return hipSuccess;
    const BinGeom g = geom_of(c);
    if (c->opt.interp == EMC_INTERP_NEAREST)
        hipLaunchKernelGGL(k_insert_binned<EMC_INTERP_NEAREST>, dim3(u1 - u0), dim3(256), 0, st, c->bin_units + 4 * u0,
                           c->bin_vals, c->bin_perm, c->bin_vox, nullptr, c->U, s0, c->T, g, c->Mnum, c->Wt);
    else
        hipLaunchKernelGGL(k_insert_binned<EMC_INTERP_TRILINEAR>, dim3(u1 - u0), dim3(256), 0, st,
                           c->bin_units + 4 * u0, c->bin_vals, c->bin_perm, c->bin_vox, c->bin_frac, c->U, s0, c->T, g,
                           c->Mnum, c->Wt);
    return hipGetLastError();
}

}  // namespace emc
