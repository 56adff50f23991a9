// C ABI of libemc: context lifetime, host<->device transfers and the stage sequence of one EMC
// iteration (main.cpp:239-480) on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "emc_internal.h"
#include "host/host_internal.h"

const char *emc_host_last_error();


using namespace emc;

namespace {

emc_status fail(emc_ctx *c, emc_status st, const std::string &msg)
{
    if (c) c->err = msg;
    else emc_host_fail(st, msg.c_str());
    return st;
}

#define HIP_TRY(c, expr)                                                                            \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fail((c), e_ == hipErrorOutOfMemory ? EMC_E_OOM : EMC_E_HIP,                     \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                         \
    } while (0)

#define EMC_TRY(expr)                                                                               \
    do {                                                                                            \
        emc_status s_ = (expr);                                                                     \
        if (s_ != EMC_OK) return s_;                                                                \
    } while (0)

template <typename T>
hipError_t dalloc(T **p, int64_t n)
{
    *p = nullptr;
    if (n <= 0) return hipSuccess;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * (size_t)n);
    if (e != hipSuccess) return e;
    e = hipMemset(*p, 0, sizeof(T) * (size_t)n);
    // the zero fill runs on the null stream, which the context's non-blocking stream does not wait
    // for: finish it before any kernel on that stream can touch the buffer
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    return e;
}

template <typename T>
void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

void ev_start(emc_ctx *c, int stage)
{
    if (c->opt.timing) (void)hipEventRecord(c->ev[2 * stage], c->stream);
}
void ev_stop(emc_ctx *c, int stage)
{
    if (c->opt.timing) (void)hipEventRecord(c->ev[2 * stage + 1], c->stream);
}

// contraction-kernel events: kind 0 = logL, 1 = M-step; chunk index k
void gev_rec(emc_ctx *c, int kind, int k, int edge)
{
    if (c->opt.timing) (void)hipEventRecord(c->gev[4 * (size_t)k + 2 * kind + edge], c->stream);
}

float gev_sum(emc_ctx *c, int kind)
{
    float tot = 0.0f;
    const int nch = (int)(c->gev.size() / 4);
    for (int k = 0; k < nch; ++k) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, c->gev[4 * (size_t)k + 2 * kind], c->gev[4 * (size_t)k + 2 * kind + 1]) == hipSuccess)
            tot += ms;
    }
    return tot;
}

emc_status sync_scalars(emc_ctx *c)
{
    HIP_TRY(c, hipMemcpyAsync(c->scal_h, c->scal, sizeof(double) * SC_N, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return EMC_OK;
}

int n_chunks(const emc_ctx *c) { return (c->S + c->chunk - 1) / c->chunk; }

// Active (mask != 0) pixels numbered j = 0..Pa-1 in 8 x 8 detector tiles (tiles row-major, pixels
// row-major inside a tile): the 64 lanes of a wave that walk consecutive j then sample a compact
// patch of the rotated detector, whose trilinear corners share far fewer cache lines than a
// 64-pixel row's (slicing), and a model brick's pixels of one rotation form longer runs of j
// (insertion).  Every kernel reads the detector position through pix[j] or the coordinates.
emc_status rebuild_active(emc_ctx *c, const float *coords_h, const int32_t *mask_h)
{
    constexpr int kTile = 8;
    std::vector<int32_t> pix;
    const int L = c->L;
    for (int ty = 0; ty < L; ty += kTile)
        for (int tx = 0; tx < L; tx += kTile)
            for (int y = ty; y < std::min(L, ty + kTile); ++y)
                for (int x = tx; x < std::min(L, tx + kTile); ++x)
                    if (mask_h[y * L + x] != 0) pix.push_back(y * L + x);
    const int Pa = (int)pix.size();
    if (Pa == 0) return fail(c, EMC_E_INVALID, "mask has no active pixel");
    const int Pp = (int)round_up(Pa, kPadP);
    c->Pa = Pa;
    c->Pp = Pp;
    std::vector<float> axyz((size_t)3 * Pp, 0.0f);
    std::vector<int32_t> pixp(Pp, -1);
    for (int j = 0; j < Pa; ++j) {
        pixp[j] = pix[j];
        if (coords_h)
            for (int a = 0; a < 3; ++a) axyz[(size_t)a * Pp + j] = coords_h[(size_t)a * c->P + pix[j]];
    }
    HIP_TRY(c, hipMemcpy(c->pix, pixp.data(), sizeof(int32_t) * Pp, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->axyz, axyz.data(), sizeof(float) * 3 * Pp, hipMemcpyHostToDevice));
    return EMC_OK;
}

// Geometry of the binned insertion changed: rebuild once rotations, coords and mask are all set.
emc_status refresh_bins(emc_ctx *c)
{
    c->bins_valid = false;
    if (c->opt.insert != EMC_INSERT_BINNED || !c->has_rot || !c->has_coords || !c->has_mask) return EMC_OK;
    std::string err;
    hipError_t e = build_bins(c, err);
    if (e != hipSuccess) return fail(c, e == hipErrorOutOfMemory ? EMC_E_OOM : EMC_E_HIP, err);
    return EMC_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------------

extern "C" void emc_default_options(emc_options *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->diff = EMC_DIFF_TRUE_POISSON;
    o->interp = EMC_INTERP_TRILINEAR;
    o->precision = EMC_PREC_BF16X2;
    o->insert = EMC_INSERT_BINNED;
    o->sigma_start = 0.4;
    o->sigma_final = 0.12;
    o->sigma_half_life = 25;
    o->device = 0;
    o->blur_sigma = 0.0;
    o->fluence = 0.0;
    o->timing = 0;
    o->mfma_tile = 0;
    o->count_bytes = 0;
}

extern "C" const char *emc_last_error(const emc_ctx *c)
{
    if (c) return c->err.c_str();
    return emc_host_last_error();
}

extern "C" emc_status emc_device_info(int32_t device, char *name, int32_t name_len, int32_t *n_cu,
                                      int64_t *hbm_bytes)
{
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return fail(nullptr, EMC_E_HIP, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (int64_t)prop.totalGlobalMem;
    return EMC_OK;
}

extern "C" void emc_destroy(emc_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipSetDevice(c->opt.device);
    if (c->prop) {
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        proposal_free(c->prop);
        delete c->prop;
        c->prop = nullptr;
    }
    dfree(c->q); dfree(c->rotm); dfree(c->logw); dfree(c->coords); dfree(c->mask); dfree(c->axyz); dfree(c->pix);
    dfree(c->K32); dfree(c->KT32); dfree(c->K16); dfree(c->K16lo); dfree(c->KT16); dfree(c->KT16lo); dfree(c->sumK);
    dfree(c->K8); dfree(c->KT8);
    dfree(c->model); dfree(c->Mnum); dfree(c->Wt); dfree(c->tmp3); dfree(c->modelq); dfree(c->taps);
    dfree(c->sumW); dfree(c->ns); dfree(c->colsum); dfree(c->colp); dfree(c->colpart); dfree(c->colmean); dfree(c->bias); dfree(c->rscale); dfree(c->cst);
    dfree(c->A0); dfree(c->A1); dfree(c->Am); dfree(c->AWm); dfree(c->K2); dfree(c->iphi); dfree(c->Pphi); dfree(c->Lh); dfree(c->Lm); dfree(c->Ll); dfree(c->lw32);
    dfree(c->Cx); dfree(c->K2h); dfree(c->K2l); dfree(c->Aw); dfree(c->G2); dfree(c->G3); dfree(c->G4);
    dfree(c->R); dfree(c->Ph); dfree(c->Pl); dfree(c->pdirty); dfree(c->runmax); dfree(c->blk); dfree(c->lse); dfree(c->part); dfree(c->red); dfree(c->T); dfree(c->U); dfree(c->scal);
    dfree(c->bin_vals); dfree(c->bin_perm); dfree(c->bin_vbeg); dfree(c->bin_frac); dfree(c->bin_units); dfree(c->bin_blocks); dfree(c->bin_groups);
    dfree(c->bin_gbeg); dfree(c->bin_part); dfree(c->bin_gpart);
    if (c->scal_h) (void)hipHostFree(c->scal_h);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : c->gev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" emc_status emc_create(const emc_dims *d, const emc_options *o, emc_ctx **out)
{
    if (!d || !out) return fail(nullptr, EMC_E_INVALID, "emc_create: null argument");
    *out = nullptr;
    emc_options opt;
    if (o) opt = *o;
    else emc_default_options(&opt);
    if (d->model_side < 4 || d->model_side > 1024 || d->n_rotations < 1 || d->n_patterns < 1 || d->slice_chunk < 0)
        return fail(nullptr, EMC_E_INVALID, "emc_create: bad dimensions");
    if (opt.diff < 0 || opt.diff > 3 || opt.interp < 0 || opt.interp > 1 || opt.precision < 0 || opt.precision > 2 ||
        opt.insert < 0 || opt.insert > 2 || (opt.count_bytes != 0 && opt.count_bytes != 2) || opt.mfma_tile < 0 ||
        opt.mstep < 0 || opt.mstep > 2 || opt.scaling < 0 || opt.scaling > 1 ||
        opt.mfma_tile > 1 || opt.tile_rows < 0 || (opt.mfma_staging != 0 && opt.mfma_staging != 1) ||
        (opt.keep_logits != 0 && opt.keep_logits != 1))
        return fail(nullptr, EMC_E_INVALID, "emc_create: bad option value");
    if (opt.scaling && opt.diff == EMC_DIFF_TRUE_POISSON)
        return fail(nullptr, EMC_E_UNSUPPORTED, "emc_create: intensity scaling needs a chi^2 diff");
    if (opt.sigma_half_life == 0 && opt.diff != EMC_DIFF_TRUE_POISSON)
        return fail(nullptr, EMC_E_INVALID, "emc_create: sigma_half_life must be non-zero");

    emc_ctx *c = new emc_ctx();
    c->dims = *d;
    c->opt = opt;
    c->L = d->model_side;
    c->S = d->n_rotations;
    c->I = d->n_patterns;
    c->P = c->L * c->L;
    c->N3 = (int64_t)c->L * c->L * c->L;
    c->chunk = d->slice_chunk > 0 ? std::min(d->slice_chunk, c->S) : c->S;
    c->chunkp = (int)round_up(c->chunk, 256);
    c->mfma = opt.precision != EMC_PREC_FP32;
    c->la_terms = opt.precision == EMC_PREC_BF16X3 ? 3 : 2;
    c->chi2_mfma = c->mfma && opt.diff != EMC_DIFF_TRUE_POISSON;
    const int nch = (c->S + c->chunk - 1) / c->chunk;
    c->Sp = (int)round_up(std::max<int64_t>((int64_t)(nch - 1) * c->chunk + c->chunkp, c->S), kPadS);
    c->Ip = (int)round_up(c->I, kPadI);
    const int PpMax = (int)round_up(c->P, kPadP);

    emc_status st = EMC_OK;
    auto A = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && st == EMC_OK)
            st = fail(c, e == hipErrorOutOfMemory ? EMC_E_OOM : EMC_E_HIP, std::string("emc_create: ") + what + ": " +
                                                                             hipGetErrorString(e));
    };
    A(hipSetDevice(opt.device), "hipSetDevice");
    if (st == EMC_OK) A(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream");
    for (auto &e : c->ev)
        if (st == EMC_OK) A(hipEventCreate(&e), "event");
    c->gev.assign(4 * (size_t)nch, nullptr);
    for (auto &e : c->gev)
        if (st == EMC_OK) A(hipEventCreate(&e), "event");
    const int64_t S = c->S, Sp = c->Sp, Ip = c->Ip, P = c->P, N3 = c->N3, cp = c->chunkp;
    if (st == EMC_OK) A(dalloc(&c->q, 4 * S), "q");
    if (st == EMC_OK) A(dalloc(&c->rotm, 12 * Sp), "rotm");
    if (st == EMC_OK) A(dalloc(&c->logw, Sp), "logw");
    if (st == EMC_OK) A(dalloc(&c->coords, 3 * P), "coords");
    if (st == EMC_OK) A(dalloc(&c->mask, P), "mask");
    if (st == EMC_OK) A(dalloc(&c->axyz, 3 * (int64_t)PpMax), "axyz");
    if (st == EMC_OK) A(dalloc(&c->pix, PpMax), "pix");
    if (st == EMC_OK) A(dalloc(&c->sumK, Ip), "sumK");
    if (c->mfma) {
        if (st == EMC_OK) A(dalloc(&c->K16, Ip * PpMax), "K16");
        if (st == EMC_OK) A(dalloc(&c->KT16, Ip * PpMax), "KT16");
        if (st == EMC_OK) A(dalloc(&c->Lh, cp * PpMax), "Lh");
        if (st == EMC_OK) A(dalloc(&c->Ll, cp * PpMax), "Ll");
        if (st == EMC_OK && c->la_terms == 3) A(dalloc(&c->Lm, cp * PpMax), "Lm");
        if (st == EMC_OK) A(dalloc(&c->Ph, Sp * Ip), "Ph");
        if (st == EMC_OK) A(dalloc(&c->Pl, Sp * Ip), "Pl");
        if (st == EMC_OK) A(dalloc(&c->pdirty, Sp * (int64_t)p_segments(c)), "pdirty");
        if (st == EMC_OK) A(hipMemset(c->pdirty, 1, sizeof(uint8_t) * Sp * p_segments(c)), "pdirty");
        if (st == EMC_OK) A(dalloc(&c->runmax, Ip), "runmax");
        if (st == EMC_OK) A(dalloc(&c->blk, (int64_t)softmax_fused_bands(c) * Ip), "blk");
        if (st == EMC_OK) A(dalloc(&c->colsum, 2 * (int64_t)PpMax), "colsum");
        if (st == EMC_OK) A(dalloc(&c->colp, colp_len(c->S, PpMax)), "colp");
        if (st == EMC_OK) A(dalloc(&c->colpart, colpart_len(PpMax)), "colpart");
        if (st == EMC_OK) A(dalloc(&c->lw32, Sp * PpMax), "lw32");
        if (c->chi2_mfma) {
            const int64_t nops = (opt.scaling ? 4 : 2) * (int64_t)c->la_terms;
            if (st == EMC_OK) A(dalloc(&c->A0, cp * PpMax), "A0");
            if (st == EMC_OK) A(dalloc(&c->A1, cp * PpMax), "A1");
            if (st == EMC_OK) A(dalloc(&c->Cx, nops * cp * PpMax), "Cx");
            if (st == EMC_OK) A(dalloc(&c->K2h, Ip * PpMax), "K2h");
            if (st == EMC_OK) A(dalloc(&c->K2l, Ip * PpMax), "K2l");
            if (st == EMC_OK) A(dalloc(&c->Aw, cp), "Aw");
            if (st == EMC_OK) A(dalloc(&c->G2, cp * Ip), "G2");
            if (opt.scaling) {
                if (st == EMC_OK) A(dalloc(&c->G3, cp * Ip), "G3");
                if (st == EMC_OK) A(dalloc(&c->G4, cp * Ip), "G4");
                if (st == EMC_OK) A(dalloc(&c->iphi, Sp * Ip), "iphi");
                if (st == EMC_OK) A(launch_fill_f32(c->iphi, 1.0f, Sp * Ip, nullptr), "iphi");
            }
        }
    } else {
        if (st == EMC_OK) A(dalloc(&c->K32, Ip * PpMax), "K32");
        if (st == EMC_OK) A(dalloc(&c->KT32, Ip * PpMax), "KT32");
        if (st == EMC_OK) A(dalloc(&c->A0, cp * PpMax), "A0");
        if (st == EMC_OK && opt.diff != EMC_DIFF_TRUE_POISSON) A(dalloc(&c->A1, cp * PpMax), "A1");
        if (opt.scaling) {
            if (st == EMC_OK) A(dalloc(&c->Am, cp * PpMax), "Am");
            if (st == EMC_OK) A(dalloc(&c->AWm, cp * PpMax), "AWm");
            if (st == EMC_OK) A(dalloc(&c->K2, Ip * PpMax), "K2");
            if (st == EMC_OK) A(dalloc(&c->iphi, Sp * Ip), "iphi");
            if (st == EMC_OK) A(dalloc(&c->Pphi, Sp * Ip), "Pphi");
            if (st == EMC_OK) A(launch_fill_f32(c->iphi, 1.0f, Sp * Ip, nullptr), "iphi");
        }
    }
    if (st == EMC_OK) A(dalloc(&c->colmean, PpMax), "colmean");
    if (st == EMC_OK) A(dalloc(&c->model, N3), "model");
    if (st == EMC_OK) A(dalloc(&c->Mnum, N3), "Mnum");
    if (st == EMC_OK) A(dalloc(&c->Wt, N3), "Wt");
    if (st == EMC_OK) A(dalloc(&c->tmp3, N3), "tmp3");
    if (st == EMC_OK) A(dalloc(&c->modelq, N3), "modelq");
    if (st == EMC_OK) A(dalloc(&c->taps, c->L), "taps");
    if (st == EMC_OK) A(dalloc(&c->sumW, Sp), "sumW");
    if (st == EMC_OK) A(dalloc(&c->ns, Sp), "ns");
    if (st == EMC_OK) A(dalloc(&c->bias, Sp), "bias");
    if (st == EMC_OK) A(dalloc(&c->rscale, Sp), "rscale");
    if (st == EMC_OK) A(dalloc(&c->cst, Ip), "cst");
    if (st == EMC_OK) A(dalloc(&c->R, Sp * Ip), "R");
    // the responsibilities pass reads the padding columns [I, Ip) of a row beside the real ones and
    // gives them p = exp(x - inf) = 0, which needs finite x: nothing else writes them
    if (st == EMC_OK) A(hipMemset(c->R, 0, sizeof(float) * Sp * Ip), "R");
    if (st == EMC_OK) A(dalloc(&c->lse, Ip), "lse");
    c->part_len = std::max<int64_t>(emc::softmax_part_len(c->S, c->I, softmax_fused_bands(c)), 4096);
    if (st == EMC_OK) A(dalloc(&c->part, c->part_len), "part");
    if (st == EMC_OK) A(dalloc(&c->red, kRedLen), "red");
    if (st == EMC_OK) A(dalloc(&c->T, Sp), "T");
    if (st == EMC_OK && !(c->mfma && opt.insert == EMC_INSERT_FUSED)) A(dalloc(&c->U, cp * PpMax), "U");
    if (st == EMC_OK) A(dalloc(&c->scal, SC_N), "scal");
    if (st == EMC_OK) A(hipHostMalloc((void **)&c->scal_h, sizeof(double) * SC_N, hipHostMallocDefault), "scal_h");
    if (st == EMC_OK) A(hipDeviceSynchronize(), "sync");  // every null-stream fill is done
    if (st == EMC_OK && opt.blur_sigma > 0.0) {
        A(launch_blur_taps(c->L, opt.blur_sigma, c->taps, c->stream), "taps");
        if (st == EMC_OK) A(hipStreamSynchronize(c->stream), "taps");
    }
    if (st != EMC_OK) {
        std::string msg = c->err;
        emc_destroy(c);
        emc_host_fail(st, msg.c_str());
        return st;
    }
    *out = c;
    return EMC_OK;
}

// ---- inputs ----------------------------------------------------------------------------------

extern "C" emc_status emc_set_rotations(emc_ctx *c, const float *q, const float *w)
{
    if (!c || !q || !w) return fail(c, EMC_E_INVALID, "set_rotations: null argument");
    for (int s = 0; s < c->S; ++s) {
        if (!(w[s] > 0.0f) || !std::isfinite(w[s])) return fail(c, EMC_E_INVALID, "set_rotations: weights must be > 0");
        const double n2 = (double)q[4 * s] * q[4 * s] + (double)q[4 * s + 1] * q[4 * s + 1] +
                          (double)q[4 * s + 2] * q[4 * s + 2] + (double)q[4 * s + 3] * q[4 * s + 3];
        if (!(std::fabs(n2 - 1.0) < 1e-3)) return fail(c, EMC_E_INVALID, "set_rotations: quaternions must be unit");
    }
    HIP_TRY(c, hipSetDevice(c->opt.device));
    float *wd = nullptr;
    HIP_TRY(c, hipMalloc((void **)&wd, sizeof(float) * c->S));
    hipError_t e = hipMemcpy(c->q, q, sizeof(float) * 4 * c->S, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(wd, w, sizeof(float) * c->S, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_rotmat(c->q, wd, c->S, c->Sp, c->rotm, c->logw, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(wd);
    HIP_TRY(c, e);
    c->has_rot = true;
    c->has_logl = c->has_resp = c->has_lse = false;
    EMC_TRY(emc_proposal_disable(c));  // its cell partition belonged to the old rotations
    return refresh_bins(c);
}

extern "C" emc_status emc_set_coords(emc_ctx *c, const float *x, const float *y, const float *z)
{
    if (!c || !x || !y || !z) return fail(c, EMC_E_INVALID, "set_coords: null argument");
    std::vector<float> h((size_t)3 * c->P);
    std::memcpy(h.data(), x, sizeof(float) * c->P);
    std::memcpy(h.data() + c->P, y, sizeof(float) * c->P);
    std::memcpy(h.data() + 2 * (size_t)c->P, z, sizeof(float) * c->P);
    for (float v : h)
        if (!std::isfinite(v)) return fail(c, EMC_E_INVALID, "set_coords: non-finite coordinate");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    HIP_TRY(c, hipMemcpy(c->coords, h.data(), sizeof(float) * 3 * c->P, hipMemcpyHostToDevice));
    c->has_coords = true;
    if (c->has_mask) {
        std::vector<int32_t> m(c->P);
        HIP_TRY(c, hipMemcpy(m.data(), c->mask, sizeof(int32_t) * c->P, hipMemcpyDeviceToHost));
        EMC_TRY(rebuild_active(c, h.data(), m.data()));
    }
    c->has_logl = c->has_resp = c->has_lse = false;
    return refresh_bins(c);
}

extern "C" emc_status emc_set_mask(emc_ctx *c, const int32_t *mask)
{
    if (!c || !mask) return fail(c, EMC_E_INVALID, "set_mask: null argument");
    for (int p = 0; p < c->P; ++p)
        if (mask[p] != 0 && mask[p] != 1) return fail(c, EMC_E_INVALID, "set_mask: values must be 0 or 1");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    HIP_TRY(c, hipMemcpy(c->mask, mask, sizeof(int32_t) * c->P, hipMemcpyHostToDevice));
    std::vector<float> h;
    if (c->has_coords) {
        h.resize((size_t)3 * c->P);
        HIP_TRY(c, hipMemcpy(h.data(), c->coords, sizeof(float) * 3 * c->P, hipMemcpyDeviceToHost));
    }
    EMC_TRY(rebuild_active(c, c->has_coords ? h.data() : nullptr, mask));
    c->has_mask = true;
    c->has_K = false;  // compaction changed
    c->has_logl = c->has_resp = c->has_lse = false;
    EMC_TRY(emc_proposal_disable(c));  // its pattern features used the old compaction
    return refresh_bins(c);
}

static emc_status set_patterns_impl(emc_ctx *c, const float *Kf)
{
    EMC_TRY(emc_proposal_disable(c));  // its features belonged to the old patterns
    const int I = c->I, Ip = c->Ip, Pp = c->Pp, Pa = c->Pa, P = c->P;
    std::vector<int32_t> pix(Pp);
    HIP_TRY(c, hipMemcpy(pix.data(), c->pix, sizeof(int32_t) * Pp, hipMemcpyDeviceToHost));
    std::vector<double> sumK(Ip, 0.0);
    double total = 0.0, maxc = 0.0;
    bool integral = true;
    for (int i = 0; i < I; ++i)
        for (int j = 0; j < Pa; ++j) {
            const float k = Kf[(size_t)i * P + pix[j]];
            if (!(k >= 0.0f) || !std::isfinite(k)) return fail(c, EMC_E_INVALID, "set_patterns: counts must be finite and >= 0");
            sumK[i] += k;
            maxc = std::max(maxc, (double)k);
            if (k != std::floor(k)) integral = false;
        }
    for (int i = 0; i < I; ++i) total += sumK[i];
    if (c->mfma) {
        // Counts are one exact BF16 term up to 256.  Above that, and for non-integer patterns
        // (read_stride-binned means, normalize_images output: emc_helper.cpp:226-236, 389-418),
        // they are a BF16 hi + lo pair: exact for integers up to 65535, else within 2^-17 relative.
        c->k_split = maxc > 256.0 || !integral;
        std::vector<bf16_t> hi((size_t)Ip * Pp, 0), lo, hT((size_t)Ip * Pp, 0), lT;
        if (c->k_split) {
            lo.assign((size_t)Ip * Pp, 0);
            lT.assign((size_t)Ip * Pp, 0);
        }
        for (int i = 0; i < I; ++i)
            for (int j = 0; j < Pa; ++j) {
                const float k = Kf[(size_t)i * P + pix[j]];
                const bf16_t h = f2bf(k);
                hi[(size_t)i * Pp + j] = h;
                hT[(size_t)j * Ip + i] = h;
                if (c->k_split) {
                    const bf16_t l = f2bf(k - bf2f(h));
                    lo[(size_t)i * Pp + j] = l;
                    lT[(size_t)j * Ip + i] = l;
                }
            }
        if (c->k_split && !c->K16lo) {
            HIP_TRY(c, dalloc(&c->K16lo, (int64_t)Ip * round_up(P, kPadP)));
            HIP_TRY(c, dalloc(&c->KT16lo, (int64_t)Ip * round_up(P, kPadP)));
        }
        // one byte per count when exact (integers <= 255) for the 256-row kernels
        const bool u8 = c->opt.count_bytes == 0 && integral && maxc <= 255.0;
        if (u8) {
            std::vector<uint8_t> b((size_t)Ip * Pp, 0), bT((size_t)Ip * Pp, 0);
            for (int i = 0; i < I; ++i)
                for (int j = 0; j < Pa; ++j) {
                    const uint8_t k = (uint8_t)Kf[(size_t)i * P + pix[j]];
                    b[(size_t)i * Pp + j] = k;
                    bT[(size_t)j * Ip + i] = k;
                }
            if (!c->K8) {
                HIP_TRY(c, dalloc(&c->K8, (int64_t)Ip * round_up(P, kPadP)));
                HIP_TRY(c, dalloc(&c->KT8, (int64_t)Ip * round_up(P, kPadP)));
            }
            HIP_TRY(c, hipMemcpy(c->K8, b.data(), b.size(), hipMemcpyHostToDevice));
            HIP_TRY(c, hipMemcpy(c->KT8, bT.data(), bT.size(), hipMemcpyHostToDevice));
        } else {
            dfree(c->K8);
            dfree(c->KT8);
        }
        HIP_TRY(c, hipMemcpy(c->K16, hi.data(), sizeof(bf16_t) * hi.size(), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(c->KT16, hT.data(), sizeof(bf16_t) * hT.size(), hipMemcpyHostToDevice));
        if (c->chi2_mfma) {
            // K^2 as a BF16 pair hi = bf16(K^2), lo = K^2 - hi: exact for integer counts <= 255
            // (K^2 < 2^16, lo < 256), else within 2^-17 relative
            std::vector<bf16_t> h2((size_t)Ip * Pp, 0), l2((size_t)Ip * Pp, 0);
            for (int i = 0; i < I; ++i)
                for (int j = 0; j < Pa; ++j) {
                    const float k = Kf[(size_t)i * P + pix[j]], k2 = k * k;
                    const bf16_t h = f2bf(k2);
                    h2[(size_t)i * Pp + j] = h;
                    l2[(size_t)i * Pp + j] = f2bf(k2 - bf2f(h));
                }
            c->k2_split = maxc > 16.0 || !integral;
            HIP_TRY(c, hipMemcpy(c->K2h, h2.data(), sizeof(bf16_t) * h2.size(), hipMemcpyHostToDevice));
            HIP_TRY(c, hipMemcpy(c->K2l, l2.data(), sizeof(bf16_t) * l2.size(), hipMemcpyHostToDevice));
        }
        if (c->k_split) {
            HIP_TRY(c, hipMemcpy(c->K16lo, lo.data(), sizeof(bf16_t) * lo.size(), hipMemcpyHostToDevice));
            HIP_TRY(c, hipMemcpy(c->KT16lo, lT.data(), sizeof(bf16_t) * lT.size(), hipMemcpyHostToDevice));
        }
    } else {
        std::vector<float> k32((size_t)Ip * Pp, 0.0f), kt((size_t)Ip * Pp, 0.0f);
        for (int i = 0; i < I; ++i)
            for (int j = 0; j < Pa; ++j) {
                const float k = Kf[(size_t)i * P + pix[j]];
                k32[(size_t)i * Pp + j] = k;
                kt[(size_t)j * Ip + i] = k;
            }
        HIP_TRY(c, hipMemcpy(c->K32, k32.data(), sizeof(float) * k32.size(), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(c->KT32, kt.data(), sizeof(float) * kt.size(), hipMemcpyHostToDevice));
        if (c->opt.scaling) HIP_TRY(c, launch_square_counts(c, c->stream));
    }
    HIP_TRY(c, hipMemcpy(c->sumK, sumK.data(), sizeof(double) * Ip, hipMemcpyHostToDevice));
    c->sumK_total = total;
    c->max_count = (int32_t)std::min(maxc, 2147483647.0);
    c->has_K = true;
    c->has_logl = c->has_resp = c->has_lse = false;
    return EMC_OK;
}

extern "C" emc_status emc_set_patterns_f32(emc_ctx *c, const float *counts)
{
    if (!c || !counts) return fail(c, EMC_E_INVALID, "set_patterns: null argument");
    if (!c->has_mask) return fail(c, EMC_E_STATE, "set_patterns: call emc_set_mask first");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    return set_patterns_impl(c, counts);
}

extern "C" emc_status emc_set_patterns_i32(emc_ctx *c, const int32_t *counts)
{
    if (!c || !counts) return fail(c, EMC_E_INVALID, "set_patterns: null argument");
    if (!c->has_mask) return fail(c, EMC_E_STATE, "set_patterns: call emc_set_mask first");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    std::vector<float> f((size_t)c->I * c->P);
    for (size_t k = 0; k < f.size(); ++k) {
        if (counts[k] < 0) return fail(c, EMC_E_INVALID, "set_patterns: counts must be >= 0");
        f[k] = (float)counts[k];
    }
    return set_patterns_impl(c, f.data());
}

extern "C" emc_status emc_set_model(emc_ctx *c, const float *model)
{
    if (!c || !model) return fail(c, EMC_E_INVALID, "set_model: null argument");
    for (int64_t v = 0; v < c->N3; ++v)
        if (!std::isfinite(model[v])) return fail(c, EMC_E_INVALID, "set_model: non-finite voxel");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    HIP_TRY(c, hipMemcpy(c->model, model, sizeof(float) * c->N3, hipMemcpyHostToDevice));
    c->has_model = true;
    c->has_logl = c->has_resp = c->has_lse = false;
    return EMC_OK;
}

// ---- outputs ---------------------------------------------------------------------------------

extern "C" emc_status emc_get_model(emc_ctx *c, float *model, float *weight)
{
    if (!c) return fail(c, EMC_E_INVALID, "get_model: null context");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (model) HIP_TRY(c, hipMemcpy(model, c->model, sizeof(float) * c->N3, hipMemcpyDeviceToHost));
    if (weight) HIP_TRY(c, hipMemcpy(weight, c->Wt, sizeof(float) * c->N3, hipMemcpyDeviceToHost));
    return EMC_OK;
}

extern "C" emc_status emc_model_stats(emc_ctx *c, double *mx, double *avg)
{
    if (!c) return fail(c, EMC_E_INVALID, "model_stats: null context");
    if (!c->has_model) return fail(c, EMC_E_STATE, "model_stats: no model");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    HIP_TRY(c, launch_normalize(c, false, false, c->stream));
    EMC_TRY(sync_scalars(c));
    if (mx) *mx = c->scal_h[SC_MAX];
    if (avg) *avg = c->scal_h[SC_AVG];
    return EMC_OK;
}

extern "C" emc_status emc_expand(emc_ctx *c, int32_t s0, int32_t count, float *out)
{
    if (!c || !out || s0 < 0 || count < 0 || s0 + count > c->S) return fail(c, EMC_E_INVALID, "expand: bad range");
    if (!c->has_rot || !c->has_coords || !c->has_model) return fail(c, EMC_E_STATE, "expand: needs rotations, coords, model");
    if (count == 0) return EMC_OK;
    HIP_TRY(c, hipSetDevice(c->opt.device));
    float *d = nullptr;
    HIP_TRY(c, hipMalloc((void **)&d, sizeof(float) * (size_t)count * c->P));
    hipError_t e = launch_expand(c, s0, count, d, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d, sizeof(float) * (size_t)count * c->P, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    HIP_TRY(c, e);
    return EMC_OK;
}

static bool bad_block(const emc_ctx *c, int32_t s0, int32_t ns, int32_t i0, int32_t ni)
{
    return s0 < 0 || ns < 0 || i0 < 0 || ni < 0 || (int64_t)s0 + ns > c->S || (int64_t)i0 + ni > c->I;
}

extern "C" emc_status emc_get_logl_block(emc_ctx *c, int32_t s0, int32_t ns, int32_t i0, int32_t ni, float *logits)
{
    if (!c || !logits) return fail(c, EMC_E_INVALID, "get_logl: null argument");
    if (bad_block(c, s0, ns, i0, ni)) return fail(c, EMC_E_INVALID, "get_logl: block out of range");
    if (!c->has_logl) return fail(c, EMC_E_STATE, "get_logl: logits not available (run emc_logl; the FP32 path "
                                                  "overwrites them in emc_normalize_resp; emc_iterate keeps them "
                                                  "all only with options.keep_logits = 1)");
    if (ns == 0 || ni == 0) return EMC_OK;
    HIP_TRY(c, hipSetDevice(c->opt.device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::vector<double> cst(ni);
    HIP_TRY(c, hipMemcpy(cst.data(), c->cst + i0, sizeof(double) * ni, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy2D(logits, sizeof(float) * ni, c->R + (size_t)s0 * c->Ip + i0, sizeof(float) * c->Ip,
                           sizeof(float) * ni, ns, hipMemcpyDeviceToHost));
    for (int s = 0; s < ns; ++s)
        for (int i = 0; i < ni; ++i) {
            float &v = logits[(size_t)s * ni + i];
            v = (float)((double)v + cst[i]);
        }
    return EMC_OK;
}

extern "C" emc_status emc_get_logl(emc_ctx *c, float *logits)
{
    if (!c) return fail(c, EMC_E_INVALID, "get_logl: null argument");
    return emc_get_logl_block(c, 0, c->S, 0, c->I, logits);
}

extern "C" emc_status emc_get_resp_block(emc_ctx *c, int32_t s0, int32_t ns, int32_t i0, int32_t ni, float *resp)
{
    if (!c || !resp) return fail(c, EMC_E_INVALID, "get_resp: null argument");
    if (bad_block(c, s0, ns, i0, ni)) return fail(c, EMC_E_INVALID, "get_resp: block out of range");
    if (!c->has_resp) return fail(c, EMC_E_STATE, "get_resp: no responsibilities");
    if (ns == 0 || ni == 0) return EMC_OK;
    HIP_TRY(c, hipSetDevice(c->opt.device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const size_t off = (size_t)s0 * c->Ip + i0;
    if (!c->mfma) {
        HIP_TRY(c, hipMemcpy2D(resp, sizeof(float) * ni, c->R + off, sizeof(float) * c->Ip, sizeof(float) * ni, ns,
                               hipMemcpyDeviceToHost));
        return EMC_OK;
    }
    std::vector<bf16_t> h((size_t)ns * ni), l((size_t)ns * ni);
    HIP_TRY(c, hipMemcpy2D(h.data(), sizeof(bf16_t) * ni, c->Ph + off, sizeof(bf16_t) * c->Ip, sizeof(bf16_t) * ni,
                           ns, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy2D(l.data(), sizeof(bf16_t) * ni, c->Pl + off, sizeof(bf16_t) * c->Ip, sizeof(bf16_t) * ni,
                           ns, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < h.size(); ++k) resp[k] = bf2f(h[k]) + bf2f(l[k]);
    if (c->opt.scaling) {  // the pair holds the M-step's P / phi: undo the factor
        std::vector<float> ip((size_t)ns * ni);
        HIP_TRY(c, hipMemcpy2D(ip.data(), sizeof(float) * ni, c->iphi + off, sizeof(float) * c->Ip, sizeof(float) * ni,
                               ns, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < ip.size(); ++k) resp[k] /= ip[k];
    }
    return EMC_OK;
}

extern "C" emc_status emc_get_resp(emc_ctx *c, float *resp)
{
    if (!c) return fail(c, EMC_E_INVALID, "get_resp: null argument");
    return emc_get_resp_block(c, 0, c->S, 0, c->I, resp);
}

extern "C" emc_status emc_get_lse(emc_ctx *c, double *lse)
{
    if (!c || !lse) return fail(c, EMC_E_INVALID, "get_lse: null argument");
    if (!c->has_lse) return fail(c, EMC_E_STATE, "get_lse: no E-step since the inputs changed");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::vector<double> cst(c->I);
    HIP_TRY(c, hipMemcpy(lse, c->lse, sizeof(double) * c->I, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(cst.data(), c->cst, sizeof(double) * c->I, hipMemcpyDeviceToHost));
    for (int i = 0; i < c->I; ++i) lse[i] += cst[i];
    return EMC_OK;
}

extern "C" emc_status emc_get_slice_weights(emc_ctx *c, double *T)
{
    if (!c || !T) return fail(c, EMC_E_INVALID, "get_slice_weights: null argument");
    if (!c->has_resp && !c->has_lse) return fail(c, EMC_E_STATE, "get_slice_weights: no responsibilities");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(T, c->T, sizeof(double) * c->S, hipMemcpyDeviceToHost));
    return EMC_OK;
}

// ---- stages ----------------------------------------------------------------------------------

// Slicing statistics, fluence, per-rotation bias and per-pattern constants (all rotations).
static emc_status slice_prepare(emc_ctx *c, float sigma)
{
    hipStream_t st = c->stream;
    ev_start(c, EMC_STAGE_SLICE);
    HIP_TRY(c, launch_slice_stats(c, c->mfma, st));
    HIP_TRY(c, launch_bias(c, sigma, st));
    HIP_TRY(c, launch_pattern_const(c, c->mfma && !c->chi2_mfma, st));
    ev_stop(c, EMC_STAGE_SLICE);
    return EMC_OK;
}

static emc_status logl_enqueue(emc_ctx *c, float sigma)
{
    hipStream_t st = c->stream;
    EMC_TRY(slice_prepare(c, sigma));
    ev_start(c, EMC_STAGE_LOGL);
    for (int s0 = 0; s0 < c->S; s0 += c->chunk) {
        const int cs = std::min(c->chunk, c->S - s0);
        const int k = s0 / c->chunk;
        if (c->chi2_mfma) {
            HIP_TRY(c, launch_slice_f32(c, s0, cs, st));
            gev_rec(c, 0, k, 0);
            HIP_TRY(c, launch_logl_chi2_mfma(c, s0, cs, st));
            gev_rec(c, 0, k, 1);
        } else if (c->mfma) {
            HIP_TRY(c, launch_slice_bf16(c, s0, cs, st));
            gev_rec(c, 0, k, 0);
            HIP_TRY(c, launch_logl_mfma(c, s0, cs, st));
            gev_rec(c, 0, k, 1);
        } else {
            HIP_TRY(c, launch_slice_f32(c, s0, cs, st));
            if (c->opt.scaling) HIP_TRY(c, launch_scaling_f32(c, s0, cs, st));
            gev_rec(c, 0, k, 0);
            HIP_TRY(c, launch_logl_f32(c, s0, cs, st));
            gev_rec(c, 0, k, 1);
        }
    }
    ev_stop(c, EMC_STAGE_LOGL);
    return EMC_OK;
}

extern "C" emc_status emc_logl(emc_ctx *c, float sigma)
{
    if (!c) return fail(c, EMC_E_INVALID, "logl: null context");
    if (!c->has_rot || !c->has_coords || !c->has_mask || !c->has_K || !c->has_model)
        return fail(c, EMC_E_STATE, "logl: needs rotations, coords, mask, patterns and model");
    if (c->opt.diff != EMC_DIFF_TRUE_POISSON && !(sigma > 0.0f)) return fail(c, EMC_E_INVALID, "logl: sigma must be > 0");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    HIP_TRY(c, launch_fill_f64(c->scal + SC_NONFINITE, 0.0, 1, c->stream));
    c->skip_logits = c->logits_partial = false;  // the caller may read every logit
    EMC_TRY(logl_enqueue(c, sigma));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->has_logl = true;
    c->has_resp = c->has_lse = false;
    return EMC_OK;
}

static emc_status softmax_enqueue(emc_ctx *c)
{
    ev_start(c, EMC_STAGE_SOFTMAX);
    HIP_TRY(c, launch_softmax(c, c->stream));
    ev_stop(c, EMC_STAGE_SOFTMAX);
    return EMC_OK;
}

static emc_status total_logl_host(emc_ctx *c, double *total)
{
    // sum_i lse_i = sum_i lse'_i + sum_i cst_i (FP64)
    std::vector<double> cst(c->I);
    HIP_TRY(c, hipMemcpy(cst.data(), c->cst, sizeof(double) * c->I, hipMemcpyDeviceToHost));
    double t = c->scal_h[SC_TOTAL];
    for (double v : cst) t += v;
    *total = t;
    return EMC_OK;
}

extern "C" emc_status emc_normalize_resp(emc_ctx *c, double *total_logl)
{
    if (!c) return fail(c, EMC_E_INVALID, "normalize_resp: null context");
    if (!c->has_logl) return fail(c, EMC_E_STATE, "normalize_resp: run emc_logl first");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    EMC_TRY(softmax_enqueue(c));
    EMC_TRY(sync_scalars(c));
    c->has_resp = c->has_lse = true;
    if (!c->mfma) c->has_logl = false;  // P overwrote the logits
    if (c->scal_h[SC_NONFINITE] != 0.0) return fail(c, EMC_E_NONFINITE, "normalize_resp: non-finite log-sum-exp");
    if (total_logl) EMC_TRY(total_logl_host(c, total_logl));
    return EMC_OK;
}

extern "C" emc_status emc_set_resp(emc_ctx *c, const float *resp)
{
    if (!c || !resp) return fail(c, EMC_E_INVALID, "set_resp: null argument");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    const int S = c->S, I = c->I;
    std::vector<double> T(c->Sp, 0.0);
    for (int s = 0; s < S; ++s) {
        double t = 0.0;
        for (int i = 0; i < I; ++i) {
            const float p = resp[(size_t)s * I + i];
            if (!(p >= 0.0f) || !std::isfinite(p)) return fail(c, EMC_E_INVALID, "set_resp: values must be finite and >= 0");
            t += p;
        }
        T[s] = t;
    }
    if (!c->mfma) {
        HIP_TRY(c, hipMemcpy2D(c->R, sizeof(float) * c->Ip, resp, sizeof(float) * I, sizeof(float) * I, S,
                               hipMemcpyHostToDevice));
        c->has_logl = false;
    } else {
        std::vector<bf16_t> h((size_t)S * I), l((size_t)S * I);
        std::vector<float> ip;
        if (c->opt.scaling) {  // the pair holds P / phi (phi of the last likelihood pass, 1 before any)
            ip.resize((size_t)S * I);
            HIP_TRY(c, hipMemcpy2D(ip.data(), sizeof(float) * I, c->iphi, sizeof(float) * c->Ip, sizeof(float) * I, S,
                                   hipMemcpyDeviceToHost));
        }
        double nnz = 0.0;  // read by the M-step's choice of form, as after an E-step
        for (size_t k = 0; k < h.size(); ++k) {
            const float q = ip.empty() ? resp[k] : resp[k] * ip[k];
            h[k] = f2bf(q);
            l[k] = f2bf(q - bf2f(h[k]));
            nnz += (h[k] | l[k]) != 0;
        }
        HIP_TRY(c, hipMemcpy(c->scal + SC_NNZ, &nnz, sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy2D(c->Ph, sizeof(bf16_t) * c->Ip, h.data(), sizeof(bf16_t) * I, sizeof(bf16_t) * I, S,
                               hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy2D(c->Pl, sizeof(bf16_t) * c->Ip, l.data(), sizeof(bf16_t) * I, sizeof(bf16_t) * I, S,
                               hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemsetAsync(c->pdirty, 1, sizeof(uint8_t) * c->Sp * p_segments(c), c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    HIP_TRY(c, hipMemcpy(c->T, T.data(), sizeof(double) * c->Sp, hipMemcpyHostToDevice));
    c->has_resp = true;
    return EMC_OK;
}

extern "C" emc_status emc_set_mstep_sums(emc_ctx *c, const float *Mnum, const float *Wt)
{
    if (!c || !Mnum || !Wt) return fail(c, EMC_E_INVALID, "set_mstep_sums: null argument");
    for (int64_t v = 0; v < c->N3; ++v)
        if (!std::isfinite(Mnum[v]) || !std::isfinite(Wt[v])) return fail(c, EMC_E_INVALID, "set_mstep_sums: non-finite value");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    HIP_TRY(c, hipMemcpy(c->Mnum, Mnum, sizeof(float) * c->N3, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->Wt, Wt, sizeof(float) * c->N3, hipMemcpyHostToDevice));
    c->has_mstep = true;
    return EMC_OK;
}

static emc_status mstep_enqueue(emc_ctx *c)
{
    hipStream_t st = c->stream;
    ev_start(c, EMC_STAGE_MSTEP);
    if (c->opt.insert != EMC_INSERT_BINNED) {  // the binned insertion's first chunk writes the sums
        HIP_TRY(c, launch_fill_f32(c->Mnum, 0.0f, c->N3, st));
        HIP_TRY(c, launch_fill_f32(c->Wt, 0.0f, c->N3, st));
    }
    const bool fused = c->mfma && c->opt.insert == EMC_INSERT_FUSED;
    // MFMA M-step form: both kernels are enqueued in AUTO mode and the one not chosen by the
    // E-step's non-zero count (scal[SC_NNZ]) returns at once, so the host never waits for it
    const int mode = fused || c->Pp > kSparseMaxPp ? EMC_MSTEP_DENSE : c->opt.mstep;
    const double thr = 0.02 * (double)c->S * (double)c->I;
    const double *gate = c->scal + SC_NNZ;
    if (!c->mfma && c->opt.scaling) HIP_TRY(c, launch_pscale_f32(c, st));  // the M-step sums K P / phi
    for (int s0 = 0; s0 < c->S; s0 += c->chunk) {
        const int cs = std::min(c->chunk, c->S - s0);
        const int k = s0 / c->chunk;
        gev_rec(c, 1, k, 0);
        if (c->mfma) {
            if (mode != EMC_MSTEP_DENSE)
                HIP_TRY(c, launch_mstep_sparse(c, s0, cs, mode == EMC_MSTEP_AUTO ? gate : nullptr, thr, st));
            if (mode != EMC_MSTEP_SPARSE)
                HIP_TRY(c, launch_mstep_mfma(c, s0, cs, fused, mode == EMC_MSTEP_AUTO ? gate : nullptr, thr, st));
        } else {
            HIP_TRY(c, launch_mstep_f32(c, s0, cs, st));
        }
        gev_rec(c, 1, k, 1);
        if (c->opt.insert == EMC_INSERT_BINNED) HIP_TRY(c, launch_insert_binned(c, s0, cs, st));
        else if (!fused) HIP_TRY(c, launch_insert(c, s0, cs, st));
    }
    ev_stop(c, EMC_STAGE_MSTEP);
    return EMC_OK;
}

extern "C" emc_status emc_maximize(emc_ctx *c)
{
    if (!c) return fail(c, EMC_E_INVALID, "maximize: null context");
    if (!c->has_resp || !c->has_K || !c->has_rot || !c->has_mask) return fail(c, EMC_E_STATE, "maximize: needs responsibilities");
    if (c->opt.insert == EMC_INSERT_BINNED && !c->bins_valid) return fail(c, EMC_E_STATE, "maximize: insertion bins not built (needs coords)");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    EMC_TRY(mstep_enqueue(c));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->has_mstep = true;
    return EMC_OK;
}

static emc_status compress_enqueue(emc_ctx *c)
{
    ev_start(c, EMC_STAGE_COMPRESS);
    // divide, normalise over the voxels that received weight, blur (the numerator, free after the
    // divide, is the blur's second buffer)
    HIP_TRY(c, launch_compress_buffers(c->L, c->opt.blur_sigma > 0.0, c->Mnum, c->Wt, c->model, c->taps, c->tmp3,
                                       c->Mnum, c->red, c->scal, c->stream));
    ev_stop(c, EMC_STAGE_COMPRESS);
    return EMC_OK;
}

extern "C" emc_status emc_compress(emc_ctx *c, double *model_avg)
{
    if (!c) return fail(c, EMC_E_INVALID, "compress: null context");
    if (!c->has_mstep) return fail(c, EMC_E_STATE, "compress: run emc_maximize first");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    HIP_TRY(c, launch_fill_f64(c->scal + SC_NONFINITE, 0.0, 1, c->stream));
    EMC_TRY(compress_enqueue(c));
    EMC_TRY(sync_scalars(c));
    if (model_avg) *model_avg = c->scal_h[SC_AVG];
    c->has_logl = c->has_resp = c->has_lse = false;
    if (c->scal_h[SC_NONFINITE] != 0.0) return fail(c, EMC_E_NONFINITE, "compress: the model is zero wherever it has weight");
    return EMC_OK;
}

extern "C" emc_status emc_normalize_model(emc_ctx *c, double *model_avg)
{
    if (!c) return fail(c, EMC_E_INVALID, "normalize_model: null context");
    if (!c->has_model) return fail(c, EMC_E_STATE, "normalize_model: no model");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    HIP_TRY(c, launch_fill_f64(c->scal + SC_NONFINITE, 0.0, 1, c->stream));
    HIP_TRY(c, launch_normalize(c, true, false, c->stream));
    EMC_TRY(sync_scalars(c));
    if (model_avg) *model_avg = c->scal_h[SC_AVG];
    c->has_logl = c->has_resp = c->has_lse = false;
    if (c->scal_h[SC_NONFINITE] != 0.0) return fail(c, EMC_E_NONFINITE, "normalize_model: no positive voxel");
    return EMC_OK;
}

// ---- proposal iteration stages (the dense stages restricted to each pattern's selected cells) --

// The grouped contraction; the slices were enqueued before the work lists were built.
static emc_status sparse_logl_enqueue(emc_ctx *c)
{
    hipStream_t st = c->stream;
    HIP_TRY(c, launch_grouped_logl(c, st));
    gev_rec(c, 0, 0, 1);
    ev_stop(c, EMC_STAGE_LOGL);
    return EMC_OK;
}

static emc_status sparse_softmax_enqueue(emc_ctx *c)
{
    ev_start(c, EMC_STAGE_SOFTMAX);
    HIP_TRY(c, launch_grouped_softmax(c, c->stream));
    ev_stop(c, EMC_STAGE_SOFTMAX);
    return EMC_OK;
}

static emc_status sparse_mstep_enqueue(emc_ctx *c)
{
    hipStream_t st = c->stream;
    ev_start(c, EMC_STAGE_MSTEP);
    if (c->opt.insert != EMC_INSERT_BINNED) {  // the binned insertion writes the sums
        HIP_TRY(c, launch_fill_f32(c->Mnum, 0.0f, c->N3, st));
        HIP_TRY(c, launch_fill_f32(c->Wt, 0.0f, c->N3, st));
    }
    gev_rec(c, 1, 0, 0);
    HIP_TRY(c, launch_grouped_mstep(c, st));
    gev_rec(c, 1, 0, 1);
    if (c->opt.insert == EMC_INSERT_BINNED) HIP_TRY(c, launch_insert_binned(c, 0, c->S, st));
    else HIP_TRY(c, launch_insert(c, 0, c->S, st));
    ev_stop(c, EMC_STAGE_MSTEP);
    return EMC_OK;
}

extern "C" emc_status emc_iterate(emc_ctx *c, int32_t iteration, emc_iter_stats *stats)
{
    if (!c) return fail(c, EMC_E_INVALID, "iterate: null context");
    if (!c->has_rot || !c->has_coords || !c->has_mask || !c->has_K || !c->has_model)
        return fail(c, EMC_E_STATE, "iterate: needs rotations, coords, mask, patterns and model");
    if (c->opt.insert == EMC_INSERT_BINNED && !c->bins_valid) return fail(c, EMC_E_STATE, "iterate: insertion bins not built");
    HIP_TRY(c, hipSetDevice(c->opt.device));
    const double sigma = emc_sigma_schedule(iteration, c->opt.sigma_start, c->opt.sigma_final, c->opt.sigma_half_life);
    if (c->opt.timing) (void)hipEventRecord(c->ev[2 * EMC_STAGE_TOTAL], c->stream);
    HIP_TRY(c, launch_fill_f64(c->scal + SC_NONFINITE, 0.0, 1, c->stream));

    // search policy: full search until the network has trained, during the warm-up, on every
    // full_every-th iteration after it, and when too many patterns would fall back
    Proposal *pr = c->prop;
    bool full = true, predicted = false, sliced = false;
    double captured = -1.0, train_loss = -1.0, pmass_mean = 0.0;
    int n_fb = 0;
    if (pr) {
        const int t = pr->iter, w = pr->o.warmup_full;
        full = pr->passes == 0 || t < w || (t - w + 1) % pr->o.full_every == 0;
        if (pr->passes > 0) {
            ev_start(c, EMC_STAGE_PROPOSE);
            EMC_TRY(proposal_predict_enqueue(c));
            ev_stop(c, EMC_STAGE_PROPOSE);
            if (!full) {
                // the slices and log-slice operands do not depend on the prediction: the device
                // computes them while the host builds the work lists
                EMC_TRY(slice_prepare(c, (float)sigma));
                ev_start(c, EMC_STAGE_LOGL);
                HIP_TRY(c, launch_slice_bf16(c, 0, c->S, c->stream));
                gev_rec(c, 0, 0, 0);
                sliced = true;
            }
            EMC_TRY(proposal_predict_wait(c));
            predicted = true;
            if (!full) {
                bool too_many = false;
                EMC_TRY(proposal_build_lists(c, &too_many));
                if (too_many) full = true;
            }
            double s = 0.0;
            for (int i = 0; i < c->I; ++i) {
                s += pr->pmass_h[i];
                n_fb += !(pr->pmass_h[i] >= pr->o.min_mass);
            }
            pmass_mean = s / c->I;
        }
    }
    if (full) {
        // negligible bands' logit stores are left out unless the caller keeps the logits or the
        // proposal's recall reads them
        c->skip_logits = !c->opt.keep_logits && !(pr && pr->o.recall > 0);
        c->logits_partial = logl_skips_stores(c);
        if (sliced) {  // one chunk covers every rotation while the proposal is enabled
            HIP_TRY(c, launch_logl_mfma(c, 0, c->S, c->stream));
            gev_rec(c, 0, 0, 1);
            ev_stop(c, EMC_STAGE_LOGL);
        } else {
            EMC_TRY(logl_enqueue(c, (float)sigma));
        }
        c->skip_logits = false;
        EMC_TRY(softmax_enqueue(c));
        EMC_TRY(mstep_enqueue(c));
        if (pr) {
            // cell sums of the full P and each cell's largest logit, then the recall from those
            // (after the capture is measured with the old recall)
            EMC_TRY(proposal_targets(c, predicted, true, &captured));
            HIP_TRY(c, launch_recall_dense(c, c->stream));
        }
    } else {
        EMC_TRY(sparse_logl_enqueue(c));
        EMC_TRY(sparse_softmax_enqueue(c));
        HIP_TRY(c, launch_recall_grouped(c, c->stream));
        EMC_TRY(sparse_mstep_enqueue(c));
    }
    EMC_TRY(compress_enqueue(c));
    if (pr) {
        // after compress, so the model-sliced training patterns come from the model the next
        // iteration uses: full iterations train on the targets plus those, proposal iterations
        // refresh on those alone
        ev_start(c, EMC_STAGE_TRAIN);
        // after the warm-up, a full search whose measured capture is at train_below or above
        // leaves the network as it is (the settled regime needs no retraining)
        const bool retrain = pr->iter < pr->o.warmup_full || !(captured >= 0.0) || pr->o.train_below >= 1.0f ||
                             captured < (double)pr->o.train_below;
        if (full) {
            if (retrain) EMC_TRY(proposal_train(c, pr->o.epochs, pr->o.train_patterns != 0, 1.0, &train_loss));
        } else if (pr->o.refresh_epochs > 0)
            EMC_TRY(proposal_train(c, pr->o.refresh_epochs, false, pr->o.refresh_fraction, &train_loss));
        ev_stop(c, EMC_STAGE_TRAIN);
    }
    if (c->opt.timing) (void)hipEventRecord(c->ev[2 * EMC_STAGE_TOTAL + 1], c->stream);
    EMC_TRY(sync_scalars(c));
    if (pr) ++pr->iter;
    c->has_logl = full && c->mfma && !c->logits_partial;
    c->has_resp = full;  // a proposal iteration leaves per-cell blocks, not the [S][I] matrix
    c->has_lse = true;   // lse and the slice sums T of either search
    c->has_mstep = true;
    if (c->scal_h[SC_NONFINITE] != 0.0) return fail(c, EMC_E_NONFINITE, "iterate: non-finite log-sum-exp or model normaliser");
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->sigma = sigma;
        stats->fluence = c->scal_h[SC_PHI0];
        stats->model_avg = c->scal_h[SC_AVG];
        stats->weight_sum = c->scal_h[SC_WSUM];
        EMC_TRY(total_logl_host(c, &stats->total_logl));
        stats->search = full ? EMC_SEARCH_FULL : EMC_SEARCH_PROPOSAL;
        stats->n_fallback = predicted ? n_fb : 0;
        stats->predicted_mass = pmass_mean;
        stats->captured_mass = captured;
        stats->train_loss = train_loss;
        stats->work_fraction = full ? 1.0 : pr->work_fraction;
        if (full && c->logits_partial)
            stats->logits_skipped = c->scal_h[SC_SKIPPED] / ((double)softmax_fused_bands(c) * c->I);
        if (c->opt.timing) {
            const int stages[] = {EMC_STAGE_SLICE, EMC_STAGE_LOGL, EMC_STAGE_SOFTMAX, EMC_STAGE_MSTEP, EMC_STAGE_COMPRESS,
                                  EMC_STAGE_TOTAL};
            for (int s : stages) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, c->ev[2 * s], c->ev[2 * s + 1]) == hipSuccess) stats->stage_ms[s] = ms;
            }
            stats->stage_ms[EMC_STAGE_LOGL_GEMM] = gev_sum(c, 0);
            stats->stage_ms[EMC_STAGE_MSTEP_GEMM] = gev_sum(c, 1);
            if (predicted) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, c->ev[2 * EMC_STAGE_PROPOSE], c->ev[2 * EMC_STAGE_PROPOSE + 1]) == hipSuccess)
                    stats->stage_ms[EMC_STAGE_PROPOSE] = ms;
            }
            if (pr) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, c->ev[2 * EMC_STAGE_TRAIN], c->ev[2 * EMC_STAGE_TRAIN + 1]) == hipSuccess)
                    stats->stage_ms[EMC_STAGE_TRAIN] = ms;
            }
        }
    }
    return EMC_OK;
}
