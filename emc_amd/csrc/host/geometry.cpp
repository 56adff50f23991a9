// Detector geometry, masks and the sigma schedule (host side).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "host_internal.h"

namespace {
std::mutex g_err_mu;
std::string g_err;
}  // namespace

emc_status emc_host_fail(emc_status st, const char *msg)
{
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = msg;
    return st;
}

std::string emc_host_last_error_str()
{
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_err;
}

const char *emc_host_last_error()
{
    std::lock_guard<std::mutex> lk(g_err_mu);
    static thread_local std::string copy;
    copy = g_err;
    return copy.c_str();
}

void emc_rotmat_host(const float q[4], float m[9])
{
    const float q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    m[0] = std::fmaf(q0, q0, std::fmaf(q1, q1, std::fmaf(-q2, q2, -(q3 * q3))));
    m[1] = 2.0f * std::fmaf(q1, q2, -(q0 * q3));
    m[2] = 2.0f * std::fmaf(q1, q3, q0 * q2);
    m[3] = 2.0f * std::fmaf(q1, q2, q0 * q3);
    m[4] = std::fmaf(q0, q0, std::fmaf(-q1, q1, std::fmaf(q2, q2, -(q3 * q3))));
    m[5] = 2.0f * std::fmaf(q2, q3, -(q0 * q1));
    m[6] = 2.0f * std::fmaf(q1, q3, -(q0 * q2));
    m[7] = 2.0f * std::fmaf(q2, q3, q0 * q1);
    m[8] = std::fmaf(q0, q0, std::fmaf(-q1, q1, std::fmaf(-q2, q2, q3 * q3)));
}

// emc_helper::calculate_coordinates (emc_helper.cpp:55-79): `real` is float there; the libm
// calls run in double and are stored to float at the same points.
extern "C" emc_status emc_coords_ewald(int32_t L, float pixel_size, float detector_distance,
                                       float wavelength, float *x, float *y, float *z)
{
    if (L < 2 || !x || !y || !z || !(wavelength > 0.0f)) return emc_host_fail(EMC_E_INVALID, "coords_ewald: bad argument");
    for (int32_t py = 0; py < L; ++py)
        for (int32_t px = 0; px < L; ++px) {
            const double dx = (double)(float)(px - L / 2) + 0.5;
            const double dy = (double)(float)(py - L / 2) + 0.5;
            const float r_pix = (float)std::sqrt(dx * dx + dy * dy);
            const float r_real = r_pix * pixel_size;
            const float ang = (float)std::atan2((double)r_real, (double)detector_distance);
            const float f_r = (float)(std::sin((double)ang) / (double)wavelength);
            const float f_z = (float)((1. - std::cos((double)ang)) / (double)wavelength);
            const int32_t p = py * L + px;
            x[p] = (float)dx;
            y[p] = (float)dy;
            z[p] = f_z / f_r * r_pix;
        }
    return EMC_OK;
}

extern "C" emc_status emc_coords_flat(int32_t L, float *x, float *y, float *z)
{
    if (L < 2 || !x || !y || !z) return emc_host_fail(EMC_E_INVALID, "coords_flat: bad argument");
    for (int32_t py = 0; py < L; ++py)
        for (int32_t px = 0; px < L; ++px) {
            const int32_t p = py * L + px;
            x[p] = (float)((double)(px - L / 2) + 0.5);
            y[p] = (float)((double)(py - L / 2) + 0.5);
            z[p] = 0.0f;
        }
    return EMC_OK;
}

extern "C" int32_t emc_mask_annulus(int32_t L, double rb, double ro, int32_t *mask)
{
    if (L < 2 || !mask) return -1;
    int32_t n = 0;
    for (int32_t py = 0; py < L; ++py)
        for (int32_t px = 0; px < L; ++px) {
            const double dx = px - L / 2 + 0.5, dy = py - L / 2 + 0.5;
            const double r = std::sqrt(dx * dx + dy * dy);
            const int32_t m = (r >= rb && r <= ro) ? 1 : 0;
            mask[py * L + px] = m;
            n += m;
        }
    return n;
}

// main.cpp:251: sigma_final + (sigma_start - sigma_final) * exp(-it / (float)h * ln 2), stored
// to a float.
extern "C" double emc_sigma_schedule(int32_t iteration, double sigma_start, double sigma_final,
                                     int32_t half_life)
{
    if (half_life == 0) return (double)(float)sigma_final;
    return (double)(float)(sigma_final + (sigma_start - sigma_final) *
                                             std::exp(-iteration / (float)half_life * std::log(2.)));
}
