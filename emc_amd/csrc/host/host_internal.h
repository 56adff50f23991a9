// Shared helpers of the host-side C++ (no GPU code).
#pragma once
#include <cstdint>

#include "../../../include/emc.h"

// Records msg as the library-wide last error (read by emc_last_error(NULL)) and returns st.
emc_status emc_host_fail(emc_status st, const char *msg);

// Geometry spec shared with the kernels: rotation matrix with an explicit FMA sequence
// (emc_cuda.cu:113-121) and the rotated model-grid coordinate u = (L/2 - 0.5) + R x.
void emc_rotmat_host(const float q[4], float m[9]);

// xoshiro256** (Blackman & Vigna 2018) seeded through splitmix64.
struct EmcRng {
    uint64_t s[4];
    explicit EmcRng(uint64_t seed);
    uint64_t next();
    double uniform();        // [0, 1), 53 random bits
    int64_t poisson(double lam);
};

// Copy of the library-wide last error message.
#include <string>
std::string emc_host_last_error_str();
