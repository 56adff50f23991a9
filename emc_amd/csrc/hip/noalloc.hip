// The allocation-free compress (emc_compress_device): the compress of emc_compress on caller-owned
// device buffers, for callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc).  Its scratch (two blur volumes for models up to
// kMaxSide^3, the statistics partials, the device scalars and the taps) is a file-scope device
// array: one copy per device, held until the process ends; the library takes its address once per
// device with hipGetSymbolAddress and passes that address to the kernels.  Every call on a device
// shares that device's copy: calls that overlap on one device race.
#include <mutex>
#include <string>

#include "../emc_internal.h"
#include "../host/host_internal.h"

namespace emc {
namespace {

constexpr int kMaxSide = 128;
constexpr size_t kVolBytes = (size_t)kMaxSide * kMaxSide * kMaxSide * sizeof(float);
constexpr size_t kOffRed = 2 * kVolBytes;
constexpr size_t kOffScal = kOffRed + kRedLen * sizeof(double);
constexpr size_t kOffTaps = kOffScal + SC_N * sizeof(double);
constexpr size_t kScratchBytes = kOffTaps + kMaxSide * sizeof(float);  // 16 MiB + 9.3 KiB

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices] = {};
std::mutex g_mu;

// This device's copy of scratch: its address is taken on the first call with the device current.
hipError_t scratch_of_current_device(void **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

// The scalars keep the last call's values: the ones the compress flags must start at zero.
__global__ void k_clear_scalars(double *__restrict__ scal)
{
    if (threadIdx.x < SC_N) scal[threadIdx.x] = 0.0;
}

__global__ void k_copy_avg(const double *__restrict__ scal, double *__restrict__ out)
{
    if (threadIdx.x == 0) *out = scal[SC_AVG];
}

emc_status fail(emc_status st, const std::string &msg)
{
    emc_host_fail(st, msg.c_str());
    return st;
}

}  // namespace
}  // namespace emc

using namespace emc;

extern "C" emc_status emc_compress_device(int32_t L, double blur_sigma, const float *Mnum, const float *Wt,
                                          float *model, double *avg, void *stream)
{
    if (!Mnum || !Wt || !model) return fail(EMC_E_INVALID, "compress_device: null buffer");
    if (L < 4 || L > kMaxSide)
        return fail(EMC_E_INVALID, "compress_device: model side " + std::to_string(L) + " outside 4.." +
                                       std::to_string(kMaxSide) + " (the static scratch holds two " +
                                       std::to_string(kMaxSide) + "^3 volumes)");
    if (!(blur_sigma >= 0.0) || !(blur_sigma < 1e6)) return fail(EMC_E_INVALID, "compress_device: bad blur_sigma");
    void *scr = nullptr;
    hipError_t e = scratch_of_current_device(&scr);
    if (e != hipSuccess) return fail(EMC_E_HIP, std::string("compress_device: scratch address: ") + hipGetErrorString(e));
    hipStream_t st = static_cast<hipStream_t>(stream);
    unsigned char *base = static_cast<unsigned char *>(scr);
    float *t1 = reinterpret_cast<float *>(base), *t2 = reinterpret_cast<float *>(base + kVolBytes);
    double *red = reinterpret_cast<double *>(base + kOffRed), *scal = reinterpret_cast<double *>(base + kOffScal);
    float *taps = reinterpret_cast<float *>(base + kOffTaps);
    const bool blur = blur_sigma > 0.0;
    hipLaunchKernelGGL(k_clear_scalars, dim3(1), dim3(64), 0, st, scal);
    e = hipGetLastError();
    if (e == hipSuccess && blur) e = launch_blur_taps(L, blur_sigma, taps, st);
    if (e == hipSuccess) e = launch_compress_buffers(L, blur, Mnum, Wt, model, taps, t1, t2, red, scal, st);
    if (e == hipSuccess && avg) {
        hipLaunchKernelGGL(k_copy_avg, dim3(1), dim3(64), 0, st, scal, avg);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return fail(EMC_E_HIP, std::string("compress_device: ") + hipGetErrorString(e));
    return EMC_OK;
}
