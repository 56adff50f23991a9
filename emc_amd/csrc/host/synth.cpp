// Synthetic EMC workloads: Gaussian-atom object -> |DFT|^2 model -> Poisson patterns.
//
// The reference reads experimental HDF5 patterns (emc_helper.cpp:198-248); the workloads this
// engine is measured on are synthetic, so the generator is part of the host library.  All
// randomness is from seeded xoshiro256** streams, so every run (and every thread count) produces
// the same arrays.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <thread>
#include <vector>

#include "host_internal.h"

// ---- RNG -----------------------------------------------------------------------------------

static inline uint64_t splitmix64(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

EmcRng::EmcRng(uint64_t seed)
{
    uint64_t x = seed;
    for (auto &v : s) v = splitmix64(x);
}

uint64_t EmcRng::next()
{
    auto rotl = [](uint64_t v, int k) { return (v << k) | (v >> (64 - k)); };
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
}

double EmcRng::uniform() { return (double)(next() >> 11) * 0x1.0p-53; }

// Poisson deviate.  lam < 10: multiplication of uniforms (Knuth, TAOCP vol. 2, 3.4.1).
// lam >= 10: PTRS transformed rejection with squeeze (W. Hormann, "The transformed rejection
// method for generating Poisson random variables", Insurance: Math. & Econ. 12, 1993).
int64_t EmcRng::poisson(double lam)
{
    if (!(lam > 0.0)) return 0;
    if (lam < 10.0) {
        const double lim = std::exp(-lam);
        double prod = uniform();
        int64_t k = 0;
        while (prod > lim) {
            prod *= uniform();
            ++k;
        }
        return k;
    }
    const double slam = std::sqrt(lam), loglam = std::log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        const double U = uniform() - 0.5;
        const double V = uniform();
        const double us = 0.5 - std::fabs(U);
        const double kf = std::floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return (int64_t)kf;
        if (kf < 0.0 || (us < 0.013 && V > us)) continue;
        if (std::log(V) + std::log(inv_alpha) - std::log(a / (us * us) + b) <=
            -lam + kf * loglam - std::lgamma(kf + 1.0))
            return (int64_t)kf;
    }
}

namespace {

template <class F>
void parallel_for(int64_t n, F &&body)
{
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (n < 64) nt = 1;
    std::vector<std::thread> th;
    const int64_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const int64_t b = (int64_t)t * chunk, e = std::min<int64_t>(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&, b, e] {
            for (int64_t i = b; i < e; ++i) body(i);
        });
    }
    for (auto &x : th) x.join();
}

using cplx = std::complex<double>;

// In-place DFT of `n` points with stride; radix-2 when n is a power of two, direct otherwise.
void dft_line(cplx *a, int n, const std::vector<cplx> &tw_direct, std::vector<cplx> &tmp)
{
    if ((n & (n - 1)) == 0) {
        for (int i = 1, j = 0; i < n; ++i) {  // bit reversal
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) std::swap(a[i], a[j]);
        }
        for (int len = 2; len <= n; len <<= 1) {
            const double ang = -2.0 * M_PI / len;
            for (int i = 0; i < n; i += len)
                for (int k = 0; k < len / 2; ++k) {
                    const cplx w(std::cos(ang * k), std::sin(ang * k));
                    const cplx u = a[i + k], v = a[i + k + len / 2] * w;
                    a[i + k] = u + v;
                    a[i + k + len / 2] = u - v;
                }
        }
        return;
    }
    tmp.assign(n, cplx(0, 0));
    for (int k = 0; k < n; ++k) {
        cplx acc(0, 0);
        for (int x = 0; x < n; ++x) acc += a[x] * tw_direct[(size_t)((int64_t)k * x % n)];
        tmp[k] = acc;
    }
    std::copy(tmp.begin(), tmp.end(), a);
}

// Trilinear sample of a float model with the kernels' geometry spec; < 0 means outside.
double tri_sample(const float *model, int L, const float m[9], float X, float Y, float Z)
{
    const float c = 0.5f * (float)L - 0.5f;
    float u[3];
    u[0] = c + std::fmaf(m[0], X, std::fmaf(m[1], Y, m[2] * Z));
    u[1] = c + std::fmaf(m[3], X, std::fmaf(m[4], Y, m[5] * Z));
    u[2] = c + std::fmaf(m[6], X, std::fmaf(m[7], Y, m[8] * Z));
    int b[3];
    double f[3];
    for (int a = 0; a < 3; ++a) {
        if (!(u[a] >= 0.0f && u[a] <= (float)(L - 1))) return -1.0;
        int i0 = (int)std::floor(u[a]);
        if (i0 > L - 2) i0 = L - 2;
        b[a] = i0;
        f[a] = (double)(u[a] - (float)i0);
    }
    double acc = 0.0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const double v = model[((size_t)(b[2] + dz) * L + (b[1] + dy)) * L + (b[0] + dx)];
                if (v < 0.0) return -1.0;
                acc += (dx ? f[0] : 1.0 - f[0]) * (dy ? f[1] : 1.0 - f[1]) * (dz ? f[2] : 1.0 - f[2]) * v;
            }
    return acc;
}

}  // namespace

extern "C" emc_status emc_synth_object(int32_t L, int32_t n_atoms, double radius, uint64_t seed,
                                       float *object)
{
    if (L < 2 || L > 1024 || n_atoms < 0 || !(radius > 0) || !object) return emc_host_fail(EMC_E_INVALID, "synth_object: bad argument");
    EmcRng rng(seed);
    std::vector<double> pos((size_t)n_atoms * 3);
    const double c0 = L / 2;
    for (int a = 0; a < n_atoms; ++a) {
        double p[3];
        do {
            for (double &v : p) v = radius * (2.0 * rng.uniform() - 1.0);
        } while (p[0] * p[0] + p[1] * p[1] + p[2] * p[2] > radius * radius);
        for (int k = 0; k < 3; ++k) pos[3 * a + k] = c0 + p[k];
    }
    const size_t N3 = (size_t)L * L * L;
    std::vector<double> rho(N3, 0.0);
    const double cut = 6.0;  // atom sigma = 1 voxel; exp(-18) ~ 1.5e-8 beyond
    for (int a = 0; a < n_atoms; ++a) {
        const double *p = &pos[3 * a];
        const int lo[3] = {std::max(0, (int)std::floor(p[0] - cut)), std::max(0, (int)std::floor(p[1] - cut)),
                           std::max(0, (int)std::floor(p[2] - cut))};
        const int hi[3] = {std::min(L - 1, (int)std::ceil(p[0] + cut)), std::min(L - 1, (int)std::ceil(p[1] + cut)),
                           std::min(L - 1, (int)std::ceil(p[2] + cut))};
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    const double d2 = (x - p[0]) * (x - p[0]) + (y - p[1]) * (y - p[1]) + (z - p[2]) * (z - p[2]);
                    if (d2 <= cut * cut) rho[((size_t)z * L + y) * L + x] += std::exp(-0.5 * d2);
                }
    }
    for (size_t v = 0; v < N3; ++v) object[v] = (float)rho[v];
    return EMC_OK;
}

extern "C" emc_status emc_synth_model(int32_t L, int32_t n_atoms, double radius, uint64_t seed,
                                      float *model)
{
    if (!model) return emc_host_fail(EMC_E_INVALID, "synth_model: null output");
    const size_t N3 = (size_t)L * L * L;
    std::vector<float> obj(N3);
    emc_status st = emc_synth_object(L, n_atoms, radius, seed, obj.data());
    if (st != EMC_OK) return st;
    std::vector<cplx> F(N3);
    for (size_t v = 0; v < N3; ++v) F[v] = cplx(obj[v], 0.0);
    std::vector<cplx> tw(L);
    for (int k = 0; k < L; ++k) tw[k] = cplx(std::cos(-2.0 * M_PI * k / L), std::sin(-2.0 * M_PI * k / L));
    for (int axis = 0; axis < 3; ++axis) {
        const size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)L : (size_t)L * L);
        parallel_for((int64_t)L * L, [&](int64_t line) {
            const int a = (int)(line % L), b = (int)(line / L);
            size_t base;
            if (axis == 0) base = ((size_t)b * L + a) * L;
            else if (axis == 1) base = (size_t)b * L * L + a;
            else base = (size_t)b * L + a;
            std::vector<cplx> buf(L), tmp;
            for (int t = 0; t < L; ++t) buf[t] = F[base + t * stride];
            dft_line(buf.data(), L, tw, tmp);
            for (int t = 0; t < L; ++t) F[base + t * stride] = buf[t];
        });
    }
    // fftshift: frequency k lands at index (k + L/2) mod L on each axis
    double sum = 0.0;
    const int h = L / 2;
    for (int z = 0; z < L; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                const double I = std::norm(F[((size_t)z * L + y) * L + x]);
                model[((size_t)((z + h) % L) * L + (y + h) % L) * L + (x + h) % L] = (float)I;
                sum += I;
            }
    const double mean = sum / (double)N3;
    if (!(mean > 0)) return emc_host_fail(EMC_E_INVALID, "synth_model: empty object");
    for (size_t v = 0; v < N3; ++v) model[v] = (float)(model[v] / mean);
    return EMC_OK;
}

extern "C" emc_status emc_synth_patterns(const float *model, int32_t L, const float *x,
                                         const float *y, const float *z, const int32_t *mask,
                                         int32_t n_patterns, double photons, uint64_t seed,
                                         int32_t *counts, float *q_true)
{
    if (!model || !x || !y || !z || !mask || !counts || L < 2 || n_patterns < 0 || !(photons >= 0))
        return emc_host_fail(EMC_E_INVALID, "synth_patterns: bad argument");
    const int P = L * L;
    parallel_for(n_patterns, [&](int64_t i) {
        EmcRng rng(seed * 0x9E3779B97F4A7C15ull + 0xD1B54A32D192ED03ull * (uint64_t)(i + 1));
        // uniform random rotation, the formula of rotations.c:30-41
        const double r1 = rng.uniform(), r2 = rng.uniform(), r3 = rng.uniform();
        float q[4] = {(float)(std::sqrt(1 - r1) * std::sin(2.0 * M_PI * r2)),
                      (float)(std::sqrt(1 - r1) * std::cos(2.0 * M_PI * r2)),
                      (float)(std::sqrt(r1) * std::sin(2.0 * M_PI * r3)),
                      (float)(std::sqrt(r1) * std::cos(2.0 * M_PI * r3))};
        if (q_true)
            for (int k = 0; k < 4; ++k) q_true[4 * i + k] = q[k];
        float m[9];
        emc_rotmat_host(q, m);
        std::vector<double> lam(P, 0.0);
        double tot = 0.0;
        for (int p = 0; p < P; ++p) {
            if (mask[p] != 1) continue;
            const double v = tri_sample(model, L, m, x[p], y[p], z[p]);
            lam[p] = v > 0.0 ? v : 0.0;
            tot += lam[p];
        }
        const double scale = tot > 0.0 ? photons / tot : 0.0;
        int32_t *K = counts + (size_t)i * P;
        for (int p = 0; p < P; ++p) K[p] = mask[p] == 1 ? (int32_t)rng.poisson(lam[p] * scale) : 0;
    });
    return EMC_OK;
}

extern "C" emc_status emc_synth_uniform_model(int32_t L, uint64_t seed, float *model)
{
    if (L < 2 || !model) return emc_host_fail(EMC_E_INVALID, "synth_uniform_model: bad argument");
    EmcRng rng(seed);
    const size_t N3 = (size_t)L * L * L;
    for (size_t v = 0; v < N3; ++v) model[v] = (float)rng.uniform();
    return EMC_OK;
}
