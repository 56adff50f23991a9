// Model initialisation modes and pattern normalisation of the reference's driver
// (emc_helper::model_init, emc_helper.cpp:440-534; normalize_images_central_part, :389-419).
// Host code: they run once, before the first iteration, as in the reference.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "host_internal.h"

// model_input = 3: radial average of all patterns.  Pixel (x, y) has radius
// r = sqrt((x - L/2 + 0.5)^2 + (y - L/2 + 0.5)^2) and shell floor(r) when r < L/2; shells are averaged
// over every pattern's pixels (masked ones included, as the reference does); voxel radius rad gives
// radavg[floor(rad)] (1 - frac) + radavg[floor(rad) + 1] frac inside L/2 and -1 outside.  The
// reference reads radavg[L/2] past its array for the outermost shell; here that shell is 0.  Sums in
// double (the reference accumulates in float).
extern "C" emc_status emc_model_init_radial(int32_t L, const float *patterns, int32_t n_patterns, float *model)
{
    if (L < 2 || L % 2 || n_patterns < 0 || !model || (n_patterns > 0 && !patterns))
        return emc_host_fail(EMC_E_INVALID, "model_init_radial: bad argument");
    const int H = L / 2;
    std::vector<double> sum(H + 1, 0.0);
    std::vector<int64_t> cnt(H + 1, 0);
    std::vector<int> shell((size_t)L * L, -1);
    for (int y = 0; y < L; ++y)
        for (int x = 0; x < L; ++x) {
            const double r = std::sqrt(std::pow(x - L / 2.0 + 0.5, 2) + std::pow(y - L / 2.0 + 0.5, 2));
            if (r < L / 2.0) shell[(size_t)y * L + x] = (int)r;
        }
    for (int64_t i = 0; i < n_patterns; ++i) {
        const float *k = patterns + (size_t)i * L * L;
        for (size_t p = 0; p < (size_t)L * L; ++p)
            if (shell[p] >= 0) {
                sum[shell[p]] += k[p];
                ++cnt[shell[p]];
            }
    }
    std::vector<double> avg(H + 1, 0.0);
    for (int r = 0; r < H; ++r) avg[r] = cnt[r] > 0 ? sum[r] / (double)cnt[r] : 0.0;
    for (int z = 0; z < L; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                const double rad = std::sqrt(std::pow(x - L / 2.0 + 0.5, 2) + std::pow(y - L / 2.0 + 0.5, 2) +
                                             std::pow(z - L / 2.0 + 0.5, 2));
                const int r = (int)rad;
                const double f = rad - r;
                model[((size_t)z * L + y) * L + x] = rad < L / 2.0 ? (float)(avg[r] * (1.0 - f) + avg[r + 1] * f) : -1.0f;
            }
    return EMC_OK;
}

// model_input = 1: every pattern inserted once at a random orientation with weight 1, then
// divided by the weight (0 where no pixel landed).  The rotation and voxel follow the reference's
// host insert_slice (emc_helper.cpp:80-123): the rotation in double, the rotated coordinate kept in a
// float, voxel
// round(L/2 + 0.5 + r) (half away from zero; the device code uses L/2 - 0.5), 0 <= v < L, pixels
// with mask == 1.  Orientations: Shoemake's uniform quaternion (quaternion_random, rotations.c:30-41)
// from this library's seeded stream (the reference draws from GSL's taus generator).  q_out
// [n_patterns][4] receives them (may be NULL).
extern "C" emc_status emc_model_init_orientations(int32_t L, const float *patterns, int32_t n_patterns,
                                                  const int32_t *mask, const float *x, const float *y,
                                                  const float *z, uint64_t seed, float *q_out, float *model)
{
    if (L < 2 || n_patterns < 0 || !model || !mask || !x || !y || !z || (n_patterns > 0 && !patterns))
        return emc_host_fail(EMC_E_INVALID, "model_init_orientations: bad argument");
    const size_t N3 = (size_t)L * L * L, P = (size_t)L * L;
    std::vector<double> num(N3, 0.0), wt(N3, 0.0);
    EmcRng rng(seed);
    for (int64_t i = 0; i < n_patterns; ++i) {
        const double r1 = rng.uniform(), r2 = rng.uniform(), r3 = rng.uniform();
        float q[4];
        q[0] = (float)(std::sqrt(1.0 - r1) * std::sin(2.0 * M_PI * r2));
        q[1] = (float)(std::sqrt(1.0 - r1) * std::cos(2.0 * M_PI * r2));
        q[2] = (float)(std::sqrt(r1) * std::sin(2.0 * M_PI * r3));
        q[3] = (float)(std::sqrt(r1) * std::cos(2.0 * M_PI * r3));
        if (q_out) std::memcpy(q_out + 4 * i, q, sizeof q);
        const double a = q[0], b = q[1], c = q[2], d = q[3];
        const double m[9] = {a * a + b * b - c * c - d * d, 2.0 * b * c - 2.0 * a * d, 2.0 * b * d + 2.0 * a * c,
                             2.0 * b * c + 2.0 * a * d, a * a - b * b + c * c - d * d, 2.0 * c * d - 2.0 * a * b,
                             2.0 * b * d - 2.0 * a * c, 2.0 * c * d + 2.0 * a * b, a * a - b * b - c * c + d * d};
        const float *k = patterns + (size_t)i * P;
        for (size_t p = 0; p < P; ++p) {
            if (mask[p] != 1) continue;
            const double X = x[p], Y = y[p], Z = z[p];
            // the reference keeps the rotated coordinate in a float (new_x, new_y, new_z)
            const float u[3] = {(float)(m[0] * X + m[1] * Y + m[2] * Z), (float)(m[3] * X + m[4] * Y + m[5] * Z),
                                (float)(m[6] * X + m[7] * Y + m[8] * Z)};
            int v[3];
            bool ok = true;
            for (int t = 0; t < 3; ++t) {
                v[t] = (int)std::round((double)L / 2.0 + 0.5 + (double)u[t]);
                ok = ok && v[t] >= 0 && v[t] < L;
            }
            if (!ok) continue;
            const size_t o = ((size_t)v[2] * L + v[1]) * L + v[0];
            num[o] += k[p];
            wt[o] += 1.0;
        }
    }
    for (size_t o = 0; o < N3; ++o) model[o] = wt[o] > 0.0 ? (float)(num[o] / wt[o]) : 0.0f;
    return EMC_OK;
}

// model_input = 2: the model read from a file.  The reference reads a libspimage HDF5 image; this
// build reads raw float32 [L^3] in the engine's voxel order (the `emc` driver's output format) and
// keeps |value|, as sp_cabs does.
extern "C" emc_status emc_model_read_raw(const char *path, int32_t L, float *model)
{
    if (!path || L < 2 || !model) return emc_host_fail(EMC_E_INVALID, "model_read_raw: bad argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return emc_host_fail(EMC_E_INVALID, (std::string("model_read_raw: cannot open ") + path).c_str());
    const size_t N3 = (size_t)L * L * L;
    const size_t got = std::fread(model, sizeof(float), N3, f);
    const bool extra = std::fgetc(f) != EOF;
    std::fclose(f);
    if (got != N3 || extra)
        return emc_host_fail(EMC_E_INVALID, "model_read_raw: the file does not hold model_side^3 float32 values");
    for (size_t v = 0; v < N3; ++v) {
        if (!std::isfinite(model[v])) return emc_host_fail(EMC_E_INVALID, "model_read_raw: non-finite value");
        model[v] = std::fabs(model[v]);
    }
    return EMC_OK;
}

// normalize_images (main.cpp:96-97 -> normalize_images_central_part, emc_helper.cpp:389-419, whose
// central mask is 1 everywhere): each pattern times P / sum over mask == 1 of its values, so its mean
// over all P pixels of the masked sum is 1.  A pattern with a zero sum is left as it is (the reference
// multiplies by P / 0).  In place is allowed (out == patterns).
extern "C" emc_status emc_normalize_patterns(const float *patterns, int32_t n_patterns, int32_t P,
                                             const int32_t *mask, float *out)
{
    if (n_patterns < 0 || P < 1 || !mask || (n_patterns > 0 && (!patterns || !out)))
        return emc_host_fail(EMC_E_INVALID, "normalize_patterns: bad argument");
    for (int64_t i = 0; i < n_patterns; ++i) {
        const float *k = patterns + (size_t)i * P;
        float *o = out + (size_t)i * P;
        double s = 0.0;
        for (int p = 0; p < P; ++p)
            if (mask[p] == 1) s += k[p];
        const double f = s != 0.0 ? (double)P / s : 1.0;
        for (int p = 0; p < P; ++p) o[p] = (float)(k[p] * f);
    }
    return EMC_OK;
}
