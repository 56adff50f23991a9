// Learned orientation proposal, host side: coarse-cell partition of the fine rotations, network
// state and training loop, prediction, and the grouped work lists of a proposal iteration.
//
// No reference counterpart: main.cpp:239-480 scores every rotation for every pattern.  The
// kernels are in hip/mlp.hip (network, features, targets, synthetic patterns) and
// hip/grouped.hip (grouped contractions and softmax); the iteration policy is in capi.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "emc_internal.h"
#include "host/host_internal.h"

using namespace emc;

namespace {

emc_status pfail(emc_ctx *c, emc_status st, const std::string &msg)
{
    if (c) c->err = msg;
    else emc_host_fail(st, msg.c_str());
    return st;
}

#define P_HIP(c, expr)                                                                              \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return pfail((c), e_ == hipErrorOutOfMemory ? EMC_E_OOM : EMC_E_HIP,                    \
                         std::string(#expr) + ": " + hipGetErrorString(e_));                        \
    } while (0)

#define P_TRY(expr)                                                                                 \
    do {                                                                                            \
        emc_status s_ = (expr);                                                                     \
        if (s_ != EMC_OK) return s_;                                                                \
    } while (0)

template <typename T>
hipError_t palloc(T **p, int64_t n)
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
void pfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <typename T>
void phfree(T *&p)
{
    if (p) (void)hipHostFree(p);
    p = nullptr;
}

inline int64_t rup(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

constexpr int D = Proposal::D;

// ---- network passes (enqueue only) ---------------------------------------------------------

emc_status forward(emc_ctx *c, const bf16_t *X, int rows, bool keep_pre)
{
    Proposal *p = c->prop;
    hipStream_t st = c->stream;
    const int H = p->H, Cp = p->Cp;
    P_HIP(c, mlp_gemm(false, false, MLP_EPI_BIAS_GELU, X, D, p->par16 + p->oW1, D, rows, H, D, keep_pre ? p->pre1 : nullptr,
                      p->h1, H, p->par + p->ob1, nullptr, rows, H, st));
    P_HIP(c, mlp_gemm(false, false, MLP_EPI_BIAS_GELU, p->h1, H, p->par16 + p->oW2, H, rows, H, H,
                      keep_pre ? p->pre2 : nullptr, p->h2, H, p->par + p->ob2, nullptr, rows, H, st));
    P_HIP(c, mlp_gemm(false, false, MLP_EPI_BIAS, p->h2, H, p->par16 + p->oW3, H, rows, Cp, H, p->z, nullptr, Cp,
                      p->par + p->ob3, nullptr, rows, Cp, st));
    return EMC_OK;
}

// Gradients of the loss whose output delta is p->dz, for the batch in p->Xb (rows rows).
emc_status backward(emc_ctx *c, int rows)
{
    Proposal *p = c->prop;
    hipStream_t st = c->stream;
    const int H = p->H, Cp = p->Cp;
    float *g = p->grad;
    P_HIP(c, mlp_gemm(true, true, MLP_EPI_STORE, p->dz, Cp, p->h2, H, Cp, H, rows, g + p->oW3, nullptr, H, nullptr,
                      nullptr, Cp, H, st));
    P_HIP(c, mlp_colsum(p->dz, rows, Cp, Cp, g + p->ob3, st));
    P_HIP(c, mlp_gemm(false, true, MLP_EPI_GELU_GRAD, p->dz, Cp, p->par16 + p->oW3, H, rows, H, Cp, nullptr, p->d2, H,
                      nullptr, p->pre2, rows, H, st));
    P_HIP(c, mlp_gemm(true, true, MLP_EPI_STORE, p->d2, H, p->h1, H, H, H, rows, g + p->oW2, nullptr, H, nullptr,
                      nullptr, H, H, st));
    P_HIP(c, mlp_colsum(p->d2, rows, H, H, g + p->ob2, st));
    P_HIP(c, mlp_gemm(false, true, MLP_EPI_GELU_GRAD, p->d2, H, p->par16 + p->oW2, H, rows, H, H, nullptr, p->d1, H,
                      nullptr, p->pre1, rows, H, st));
    P_HIP(c, mlp_gemm(true, true, MLP_EPI_STORE, p->d1, H, p->Xb, D, H, D, rows, g + p->oW1, nullptr, D, nullptr,
                      nullptr, H, D, st));
    P_HIP(c, mlp_colsum(p->d1, rows, H, H, g + p->ob1, st));
    return EMC_OK;
}

// Loss and gradients of the batch in Xb / tb (rows_valid real rows of o.batch); loss summed
// into scal[SC_LOSS].
emc_status loss_and_grad(emc_ctx *c, int rows_valid)
{
    Proposal *p = c->prop;
    const int B = p->o.batch;
    P_TRY(forward(c, p->Xb, B, true));
    P_HIP(c, mlp_softmax_ce(p->z, p->Cp, p->C, rows_valid, B, p->tb, 1.0f / (float)rows_valid, p->prob, p->dz,
                            p->loss_part, c->stream));
    P_TRY(backward(c, B));
    P_HIP(c, mlp_accum(p->loss_part, B, c->scal + SC_LOSS, c->stream));
    return EMC_OK;
}

emc_status adam_step(emc_ctx *c)
{
    Proposal *p = c->prop;
    ++p->adam_t;
    P_HIP(c, mlp_adam(p->par, p->grad, p->am, p->av, p->par16, p->n_par, p->o.lr, p->adam_t, c->stream));
    return EMC_OK;
}

emc_status read_loss(emc_ctx *c, double *out)
{
    double v = 0.0;
    P_HIP(c, hipMemcpyAsync(&v, c->scal + SC_LOSS, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    P_HIP(c, hipStreamSynchronize(c->stream));
    *out = v;
    return EMC_OK;
}

// Test batches: x [B][D] (and t [B][C]) into Xb / tb, zero-padded to o.batch rows.
emc_status upload_batch(emc_ctx *c, const float *x, const float *t, int B)
{
    Proposal *p = c->prop;
    if (B < 1 || B > p->o.batch) return pfail(c, EMC_E_INVALID, "mlp: batch rows must be in [1, options.batch]");
    std::vector<bf16_t> xb((size_t)p->o.batch * D, 0);
    for (size_t k = 0; k < (size_t)B * D; ++k) {
        if (!std::isfinite(x[k])) return pfail(c, EMC_E_INVALID, "mlp: non-finite feature");
        xb[k] = f2bf(x[k]);
    }
    P_HIP(c, hipMemcpy(p->Xb, xb.data(), sizeof(bf16_t) * xb.size(), hipMemcpyHostToDevice));
    if (t) {
        std::vector<float> tb((size_t)p->o.batch * p->Cp, 0.0f);
        for (int b = 0; b < B; ++b)
            for (int k = 0; k < p->C; ++k) tb[(size_t)b * p->Cp + k] = t[(size_t)b * p->C + k];
        P_HIP(c, hipMemcpy(p->tb, tb.data(), sizeof(float) * tb.size(), hipMemcpyHostToDevice));
    }
    return EMC_OK;
}

// Largest packed work lists (int32 elements) and slot count any proposal iteration can need: at
// most floor(max_fallback * I) patterns take every non-empty cell (more and the iteration
// searches fully instead), the others at most top_k + recall cells; every count of tiles follows
// from those slot counts per cell.  Allocated once by emc_proposal_enable, so proposal iterations
// never allocate.
int64_t list_bound(const emc_ctx *c, const Proposal *p, int64_t *slots_max)
{
    const int I = c->I, C = p->C, Pp = c->Pp;
    int64_t cne = 0, mt_sum = 0, mt_max = 0;
    for (int cc = 0; cc < C; ++cc) {
        const int64_t nr = p->rbeg_h[cc + 1] - p->rbeg_h[cc];
        if (nr == 0) continue;
        ++cne;
        const int64_t mt = (nr + 127) / 128;
        mt_sum += mt;
        mt_max = std::max(mt_max, mt);
    }
    const int64_t nfb = std::min<int64_t>(I, (int64_t)std::floor((double)p->o.max_fallback * I));
    const int64_t per = std::min<int64_t>(p->o.top_k + p->o.recall, cne);
    const int64_t slots = nfb * cne + (I - nfb) * per;
    const int64_t n_te = mt_max * ((slots + 127) / 128 + cne);
    const int64_t n_tm = mt_sum * (Pp / 128);
    const int64_t n_ts = (slots + 63) / 64 + cne;
    auto sec = [](int64_t n) { return rup(std::max<int64_t>(n, 1), 4); };
    *slots_max = slots;
    return sec(slots) + sec(C + 1) + sec(C) + sec(I + 1) + sec(slots) + sec(4 * n_te) + sec(4 * n_tm) + sec(4 * n_ts);
}

emc_status need_prop(emc_ctx *c, const char *what)
{
    if (!c) return pfail(nullptr, EMC_E_INVALID, std::string(what) + ": null context");
    if (!c->prop) return pfail(c, EMC_E_STATE, std::string(what) + ": proposal not enabled");
    return EMC_OK;
}

}  // namespace

namespace emc {

void proposal_free(Proposal *p)
{
    if (!p) return;
    pfree(p->cell_of); pfree(p->cell_rot); pfree(p->rbeg);
    pfree(p->par); pfree(p->grad); pfree(p->am); pfree(p->av); pfree(p->par16);
    pfree(p->mu); pfree(p->istd); pfree(p->X); pfree(p->raw);
    pfree(p->pre1); pfree(p->pre2); pfree(p->z); pfree(p->prob);
    pfree(p->h1); pfree(p->h2); pfree(p->dz); pfree(p->d1); pfree(p->d2); pfree(p->loss_part);
    pfree(p->Tt); pfree(p->Mx); pfree(p->Xb); pfree(p->tb); pfree(p->bidx); pfree(p->syn_counts); pfree(p->Xs); pfree(p->syn_rot);
    pfree(p->topc); pfree(p->pmass); pfree(p->cap); pfree(p->recall);
    phfree(p->topc_h); phfree(p->pmass_h); phfree(p->recall_h);
    if (p->ev_pred) (void)hipEventDestroy(p->ev_pred);
    p->ev_pred = nullptr;
    pfree(p->lists); phfree(p->lists_h); pfree(p->boff); phfree(p->boff_h);
    pfree(p->slot_ps); pfree(p->slot_pm); pfree(p->slot_lse);
}

emc_status proposal_predict_enqueue(emc_ctx *c)
{
    Proposal *p = c->prop;
    hipStream_t st = c->stream;
    p->pred_valid = false;
    P_TRY(forward(c, p->X, c->Ip, false));
    P_HIP(c, mlp_softmax_ce(p->z, p->Cp, p->C, c->I, c->Ip, nullptr, 0.0f, p->prob, nullptr, nullptr, st));
    P_HIP(c, mlp_topk(p->prob, p->Cp, p->C, c->I, p->o.top_k, p->topc, p->pmass, st));
    P_HIP(c, mlp_union_mass(c, st));  // pmass: predicted mass of every cell the iteration would search
    P_HIP(c, hipMemcpyAsync(p->topc_h, p->topc, sizeof(int32_t) * (size_t)c->I * p->o.top_k, hipMemcpyDeviceToHost, st));
    P_HIP(c, hipMemcpyAsync(p->pmass_h, p->pmass, sizeof(float) * c->I, hipMemcpyDeviceToHost, st));
    if (p->o.recall > 0)
        P_HIP(c, hipMemcpyAsync(p->recall_h, p->recall, sizeof(int32_t) * (size_t)c->I * p->o.recall,
                                hipMemcpyDeviceToHost, st));
    P_HIP(c, hipEventRecord(p->ev_pred, st));
    return EMC_OK;
}

emc_status proposal_predict_wait(emc_ctx *c)
{
    P_HIP(c, hipEventSynchronize(c->prop->ev_pred));
    c->prop->pred_valid = true;
    return EMC_OK;
}

emc_status proposal_predict(emc_ctx *c)
{
    P_TRY(proposal_predict_enqueue(c));
    return proposal_predict_wait(c);
}

emc_status proposal_build_lists(emc_ctx *c, bool *too_many)
{
    Proposal *p = c->prop;
    const int I = c->I, C = p->C, k = p->o.top_k, R = p->o.recall, S = c->S, Pp = c->Pp;
    const std::vector<int32_t> &rbeg = p->rbeg_h;
    auto nr = [&](int cc) { return rbeg[cc + 1] - rbeg[cc]; };

    // cells searched for pattern i (unless it falls back): the predicted top k, then the
    // recalled cells not among them
    auto cells_of = [&](int i, int32_t *out) {
        int n = 0;
        for (int q = 0; q < k; ++q) out[n++] = p->topc_h[(size_t)i * k + q];
        for (int q = 0; q < R; ++q) {
            const int32_t cr = p->recall_h[(size_t)i * R + q];
            if (cr < 0 || cr >= C) continue;
            bool seen = false;
            for (int u = 0; u < n; ++u) seen |= out[u] == cr;
            if (!seen) out[n++] = cr;
        }
        return n;
    };
    std::vector<int32_t> cs(k + R);

    // patterns searched over every cell: low predicted mass, or no non-empty cell among the top k
    std::vector<uint8_t> fb(I, 0);
    int n_fb = 0;
    for (int i = 0; i < I; ++i) {
        bool any = false;
        for (int q = 0; q < k; ++q) any |= nr(p->topc_h[(size_t)i * k + q]) > 0;
        fb[i] = !(p->pmass_h[i] >= p->o.min_mass) || !any;
        n_fb += fb[i];
    }
    p->n_fallback = n_fb;
    *too_many = (double)n_fb > (double)p->o.max_fallback * I;
    if (*too_many) return EMC_OK;

    std::vector<int64_t> cnt(C + 1, 0);
    for (int i = 0; i < I; ++i) {
        if (fb[i]) {
            for (int cc = 0; cc < C; ++cc) cnt[cc] += nr(cc) > 0;
        } else {
            const int n = cells_of(i, cs.data());
            for (int q = 0; q < n; ++q) cnt[cs[q]] += nr(cs[q]) > 0;
        }
    }
    std::vector<int64_t> sbeg(C + 1, 0);
    std::vector<int32_t> nsp(C, 0);
    int64_t blk = 0;
    double pairs = 0.0;
    for (int cc = 0; cc < C; ++cc) {
        sbeg[cc + 1] = sbeg[cc] + cnt[cc];
        nsp[cc] = cnt[cc] ? (int32_t)rup(cnt[cc], 32) : 0;
        p->boff_h[cc] = blk;
        blk += (int64_t)nr(cc) * nsp[cc];
        pairs += (double)nr(cc) * (double)cnt[cc];
    }
    const int64_t n_slots = sbeg[C];
    if (n_slots > INT32_MAX / 2 || blk > (int64_t)c->Sp * c->Ip)
        return pfail(c, EMC_E_UNSUPPORTED, "proposal: work lists too large");
    p->n_slots = n_slots;
    p->n_blk = blk;
    p->work_fraction = pairs / ((double)S * (double)I);

    // tiles
    int64_t n_te = 0, n_tm = 0, n_ts = 0;
    for (int cc = 0; cc < C; ++cc) {
        if (!cnt[cc]) continue;
        const int64_t mt = (nr(cc) + 127) / 128;
        n_te += mt * ((cnt[cc] + 127) / 128);
        n_tm += mt * (Pp / 128);
        n_ts += (cnt[cc] + 63) / 64;
    }
    // packed int32 lists, each section starting on a 16-byte boundary
    int64_t off = 0;
    auto sec = [&](int64_t n) {
        const int64_t o = off;
        off += rup(std::max<int64_t>(n, 1), 4);
        return o;
    };
    p->off_slot_pat = sec(n_slots);
    p->off_sbeg = sec(C + 1);
    p->off_nsp = sec(C);
    p->off_pbeg = sec(I + 1);
    p->off_pslot = sec(n_slots);
    p->off_tile_e = sec(4 * n_te);
    p->off_tile_m = sec(4 * n_tm);
    p->off_tile_sm = sec(4 * n_ts);
    const int64_t need = off;
    if (need > p->list_cap || n_slots > p->slot_cap)  // the bound of list_bound() holds by construction
        return pfail(c, EMC_E_STATE, "proposal: work lists exceed their preallocated bound");
    int32_t *L = p->lists_h;
    int32_t *slot_pat = L + p->off_slot_pat, *pbeg = L + p->off_pbeg, *pslot = L + p->off_pslot;
    for (int cc = 0; cc <= C; ++cc) L[p->off_sbeg + cc] = (int32_t)sbeg[cc];
    for (int cc = 0; cc < C; ++cc) L[p->off_nsp + cc] = nsp[cc];
    std::vector<int64_t> cur(sbeg.begin(), sbeg.end() - 1);
    int64_t q_out = 0;
    pbeg[0] = 0;
    for (int i = 0; i < I; ++i) {
        auto take = [&](int cc) {
            if (nr(cc) == 0) return;
            const int64_t s = cur[cc]++;
            slot_pat[s] = i;
            pslot[q_out++] = (int32_t)s;
        };
        if (fb[i]) {
            for (int cc = 0; cc < C; ++cc) take(cc);
        } else {
            const int n = cells_of(i, cs.data());
            for (int q = 0; q < n; ++q) take(cs[q]);
        }
        pbeg[i + 1] = (int32_t)q_out;
    }
    int32_t *te = L + p->off_tile_e, *tm = L + p->off_tile_m, *ts = L + p->off_tile_sm;
    for (int cc = 0; cc < C; ++cc) {
        if (!cnt[cc]) continue;
        for (int m0 = 0; m0 < nr(cc); m0 += 128) {
            for (int n0 = 0; n0 < cnt[cc]; n0 += 128) {
                te[0] = cc; te[1] = m0; te[2] = n0; te[3] = 0;
                te += 4;
            }
            for (int n0 = 0; n0 < Pp; n0 += 128) {
                tm[0] = cc; tm[1] = m0; tm[2] = n0; tm[3] = 0;
                tm += 4;
            }
        }
        for (int s0 = 0; s0 < cnt[cc]; s0 += 64) {
            ts[0] = cc; ts[1] = s0; ts[2] = 0; ts[3] = 0;
            ts += 4;
        }
    }
    p->n_tiles_e = (int)n_te;
    p->n_tiles_m = (int)n_tm;
    p->n_tiles_sm = (int)n_ts;

    P_HIP(c, hipMemcpyAsync(p->lists, p->lists_h, sizeof(int32_t) * (size_t)need, hipMemcpyHostToDevice, c->stream));
    P_HIP(c, hipMemcpyAsync(p->boff, p->boff_h, sizeof(int64_t) * (size_t)C, hipMemcpyHostToDevice, c->stream));
    return EMC_OK;
}

emc_status proposal_targets(emc_ctx *c, bool measure, bool logits, double *captured)
{
    Proposal *p = c->prop;
    P_HIP(c, mlp_cell_targets(c, p->Tt, logits && p->o.recall > 0 ? p->Mx : nullptr, c->stream));
    p->targets_valid = true;
    if (captured) *captured = -1.0;
    if (measure && p->pred_valid) {
        P_HIP(c, mlp_captured(c, p->cap, c->stream));
        std::vector<double> cap(c->I);
        P_HIP(c, hipMemcpyAsync(cap.data(), p->cap, sizeof(double) * c->I, hipMemcpyDeviceToHost, c->stream));
        P_HIP(c, hipStreamSynchronize(c->stream));
        double s = 0.0;
        for (double v : cap) s += v;
        if (captured) *captured = s / c->I;
        p->cap_h = std::move(cap);
    }
    return EMC_OK;
}

emc_status proposal_train(emc_ctx *c, int epochs, bool use_patterns, double syn_fraction, double *loss)
{
    Proposal *p = c->prop;
    hipStream_t st = c->stream;
    if (use_patterns && !p->targets_valid) return pfail(c, EMC_E_STATE, "proposal_train: no responsibility targets yet");
    const int I = c->I, B = p->o.batch;
    // each pass draws syn_fraction of the synthetic patterns afresh
    const int n_syn = p->n_syn > 0 ? std::max(1, (int)std::lround(p->n_syn * syn_fraction)) : 0;
    if (!use_patterns && n_syn == 0) {
        if (loss) *loss = -1.0;
        return EMC_OK;
    }
    // training items: patterns 0..I-1 (cell targets) then synthetic patterns I..I+n_syn-1
    const int64_t N = use_patterns ? (int64_t)I + n_syn : n_syn, first = use_patterns ? 0 : I;
    EmcRng rng(p->rng_state);
    std::vector<int32_t> idx(N + B), srot(std::max(n_syn, 1));
    double last = -1.0;
    for (int e = 0; e < epochs; ++e) {
        if (n_syn > 0) {
            for (int k = 0; k < n_syn; ++k) srot[k] = (int32_t)(rng.next() % (uint64_t)c->S);
            P_HIP(c, hipMemcpyAsync(p->syn_rot, srot.data(), sizeof(int32_t) * n_syn, hipMemcpyHostToDevice, st));
            const float photons = (float)(c->sumK_total / I);
            P_HIP(c, mlp_synth_counts(c, n_syn, photons, rng.next(), st));
            const int pad = (int)rup(n_syn, 64);
            P_HIP(c, mlp_features(p->syn_counts, nullptr, nullptr, c->Pp, c->Pa, c->pix, c->L, n_syn, pad, p->raw, st));
            P_HIP(c, mlp_standardise(p->raw, n_syn, pad, p->mu, p->istd, p->Xs, st));
        }
        for (int64_t k = 0; k < N; ++k) idx[k] = (int32_t)(first + k);
        for (int64_t k = N - 1; k > 0; --k) std::swap(idx[k], idx[rng.next() % (uint64_t)(k + 1)]);
        for (int k = 0; k < B; ++k) idx[N + k] = idx[0];
        P_HIP(c, hipMemcpyAsync(p->bidx, idx.data(), sizeof(int32_t) * (size_t)(N + B), hipMemcpyHostToDevice, st));
        P_HIP(c, launch_fill_f64(c->scal + SC_LOSS, 0.0, 1, st));
        for (int64_t b0 = 0; b0 < N; b0 += B) {
            const int rows_valid = (int)std::min<int64_t>(B, N - b0);
            P_HIP(c, mlp_gather_batch(c, p->bidx + b0, B, st));
            P_TRY(loss_and_grad(c, rows_valid));
            P_TRY(adam_step(c));
        }
        ++p->passes;
        double tot = 0.0;
        P_TRY(read_loss(c, &tot));  // also keeps the host vectors alive until the copies are done
        last = tot / (double)N;
    }
    p->rng_state = rng.next();
    if (epochs > 0) p->last_loss = last;
    if (loss) *loss = last;
    return EMC_OK;
}

}  // namespace emc

// ---- C ABI -----------------------------------------------------------------------------------

extern "C" void emc_proposal_defaults(emc_proposal_options *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->top_k = 8;
    o->warmup_full = 3;
    o->full_every = 5;
    o->hidden = 512;
    o->epochs = 4;
    o->batch = 1024;  // 256 at Adam step 1e-3 trained 1.7x slower for no better capture
    o->synth_patterns = -1;
    o->refresh_epochs = 1;
    o->recall = 8;
    o->train_patterns = 1;
    o->refresh_fraction = 0.25f;
    o->train_below = 0.9999f;
    o->min_mass = 0.99f;
    o->lr = 1.5e-3f;
    o->max_fallback = 0.1f;
    o->seed = 1;
}

extern "C" emc_status emc_coarse_partition(const float *fine, int32_t n_fine, const float *coarse, int32_t n_coarse,
                                           int32_t *cell)
{
    if (!fine || !coarse || !cell || n_fine < 1 || n_coarse < 1)
        return pfail(nullptr, EMC_E_INVALID, "coarse_partition: bad argument");
    for (int32_t s = 0; s < n_fine; ++s) {
        const float *f = fine + 4 * (size_t)s;
        double best = -1.0;
        int32_t bc = 0;
        for (int32_t k = 0; k < n_coarse; ++k) {
            const float *q = coarse + 4 * (size_t)k;
            const double d = std::fabs((double)f[0] * q[0] + (double)f[1] * q[1] + (double)f[2] * q[2] + (double)f[3] * q[3]);
            if (d > best) {
                best = d;
                bc = k;
            }
        }
        cell[s] = bc;
    }
    return EMC_OK;
}

extern "C" emc_status emc_proposal_disable(emc_ctx *c)
{
    if (!c) return pfail(nullptr, EMC_E_INVALID, "proposal_disable: null context");
    if (c->prop) {
        (void)hipSetDevice(c->opt.device);
        (void)hipStreamSynchronize(c->stream);
        proposal_free(c->prop);
        delete c->prop;
        c->prop = nullptr;
    }
    return EMC_OK;
}

static emc_status enable_impl(emc_ctx *c, const float *coarse, int32_t nc, const emc_proposal_options &o)
{
    Proposal *p = c->prop;
    p->o = o;
    p->C = nc;
    p->Cp = (int)rup(nc, 64);
    p->H = o.hidden;
    const int S = c->S, I = c->I, Ip = c->Ip, H = p->H, Cp = p->Cp, B = o.batch;
    hipStream_t st = c->stream;

    // partition of the fine rotations into coarse cells, cell-major order
    std::vector<float> q((size_t)4 * S);
    P_HIP(c, hipMemcpy(q.data(), c->q, sizeof(float) * q.size(), hipMemcpyDeviceToHost));
    p->cell_h.assign(S, 0);
    P_TRY(emc_coarse_partition(q.data(), S, coarse, nc, p->cell_h.data()));
    p->rbeg_h.assign(nc + 1, 0);
    for (int s = 0; s < S; ++s) ++p->rbeg_h[p->cell_h[s] + 1];
    for (int k = 0; k < nc; ++k) p->rbeg_h[k + 1] += p->rbeg_h[k];
    p->cell_rot_h.assign(S, 0);
    {
        std::vector<int32_t> cur(p->rbeg_h.begin(), p->rbeg_h.end() - 1);
        for (int s = 0; s < S; ++s) p->cell_rot_h[cur[p->cell_h[s]]++] = s;
    }
    P_HIP(c, palloc(&p->cell_of, S));
    P_HIP(c, palloc(&p->cell_rot, S));
    P_HIP(c, palloc(&p->rbeg, nc + 1));
    P_HIP(c, hipMemcpy(p->cell_of, p->cell_h.data(), sizeof(int32_t) * S, hipMemcpyHostToDevice));
    P_HIP(c, hipMemcpy(p->cell_rot, p->cell_rot_h.data(), sizeof(int32_t) * S, hipMemcpyHostToDevice));
    P_HIP(c, hipMemcpy(p->rbeg, p->rbeg_h.data(), sizeof(int32_t) * (nc + 1), hipMemcpyHostToDevice));

    // parameters: W1 | b1 | W2 | b2 | W3 [Cp][H] | b3 [Cp]; He-uniform hidden layers, a smaller
    // output layer so the first predictions are close to uniform
    p->oW1 = 0;
    p->ob1 = (int64_t)H * D;
    p->oW2 = p->ob1 + H;
    p->ob2 = p->oW2 + (int64_t)H * H;
    p->oW3 = p->ob2 + H;
    p->ob3 = p->oW3 + (int64_t)Cp * H;
    p->n_par = p->ob3 + Cp;
    P_HIP(c, palloc(&p->par, p->n_par));
    P_HIP(c, palloc(&p->grad, p->n_par));
    P_HIP(c, palloc(&p->am, p->n_par));
    P_HIP(c, palloc(&p->av, p->n_par));
    P_HIP(c, palloc(&p->par16, p->n_par));
    {
        std::vector<float> w(p->n_par, 0.0f);
        EmcRng rng(o.seed);
        auto fill = [&](int64_t off, int64_t n, double a) {
            for (int64_t k = 0; k < n; ++k) w[off + k] = (float)((2.0 * rng.uniform() - 1.0) * a);
        };
        fill(p->oW1, (int64_t)H * D, std::sqrt(6.0 / D));
        fill(p->oW2, (int64_t)H * H, std::sqrt(6.0 / H));
        fill(p->oW3, (int64_t)nc * H, std::sqrt(1.0 / H));
        p->rng_state = rng.next();
        P_HIP(c, hipMemcpy(p->par, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
        P_HIP(c, mlp_to_bf16(p->par, p->n_par, p->par16, st));
    }

    // features of the patterns (counts are fixed for the context's lifetime)
    p->n_syn = o.synth_patterns < 0 ? I / 4 : o.synth_patterns;
    const int syn_pad = (int)rup(std::max(p->n_syn, 1), 64);
    P_HIP(c, palloc(&p->raw, (int64_t)std::max(Ip, syn_pad) * D));
    P_HIP(c, palloc(&p->mu, D));
    P_HIP(c, palloc(&p->istd, D));
    P_HIP(c, palloc(&p->X, (int64_t)Ip * D));
    P_HIP(c, mlp_features(nullptr, c->K8 ? nullptr : c->K16, c->K8, c->Pp, c->Pa, c->pix, c->L, I, Ip, p->raw, st));
    P_HIP(c, mlp_feat_stats(p->raw, I, p->mu, p->istd, st));
    P_HIP(c, mlp_standardise(p->raw, I, Ip, p->mu, p->istd, p->X, st));

    // activations and training buffers
    const int64_t rows = std::max(Ip, B);
    P_HIP(c, palloc(&p->pre1, (int64_t)B * H));
    P_HIP(c, palloc(&p->pre2, (int64_t)B * H));
    P_HIP(c, palloc(&p->h1, rows * H));
    P_HIP(c, palloc(&p->h2, rows * H));
    P_HIP(c, palloc(&p->z, rows * Cp));
    P_HIP(c, palloc(&p->prob, rows * Cp));
    P_HIP(c, palloc(&p->dz, (int64_t)B * Cp));
    P_HIP(c, palloc(&p->d1, (int64_t)B * H));
    P_HIP(c, palloc(&p->d2, (int64_t)B * H));
    P_HIP(c, palloc(&p->loss_part, B));
    P_HIP(c, palloc(&p->Tt, (int64_t)Cp * Ip));
    if (o.recall > 0) P_HIP(c, palloc(&p->Mx, (int64_t)Cp * Ip));
    P_HIP(c, palloc(&p->Xb, (int64_t)B * D));
    P_HIP(c, palloc(&p->tb, (int64_t)B * Cp));
    P_HIP(c, palloc(&p->bidx, (int64_t)I + p->n_syn + B));
    P_HIP(c, palloc(&p->syn_counts, (int64_t)syn_pad * c->Pp));
    P_HIP(c, palloc(&p->Xs, (int64_t)syn_pad * D));
    P_HIP(c, palloc(&p->syn_rot, syn_pad));
    P_HIP(c, palloc(&p->topc, (int64_t)Ip * o.top_k));
    P_HIP(c, hipHostMalloc((void **)&p->topc_h, sizeof(int32_t) * (size_t)I * o.top_k, hipHostMallocDefault));
    P_HIP(c, hipHostMalloc((void **)&p->pmass_h, sizeof(float) * (size_t)I, hipHostMallocDefault));
    P_HIP(c, hipHostMalloc((void **)&p->recall_h, sizeof(int32_t) * (size_t)I * std::max(o.recall, 1), hipHostMallocDefault));
    P_HIP(c, hipEventCreateWithFlags(&p->ev_pred, hipEventDisableTiming));
    P_HIP(c, palloc(&p->pmass, Ip));
    P_HIP(c, palloc(&p->cap, Ip));
    P_HIP(c, palloc(&p->recall, (int64_t)Ip * std::max(o.recall, 1)));
    P_HIP(c, hipMemsetAsync(p->recall, 0xff, sizeof(int32_t) * (size_t)Ip * std::max(o.recall, 1), st));  // -1: none yet
    P_HIP(c, palloc(&p->boff, nc));
    P_HIP(c, hipHostMalloc((void **)&p->boff_h, sizeof(int64_t) * nc, hipHostMallocDefault));
    // grouped work lists and their slot scratch at their largest size (iterations never allocate)
    int64_t slots_max = 0;
    p->list_cap = list_bound(c, p, &slots_max);
    p->slot_cap = std::max<int64_t>(slots_max, 1);
    P_HIP(c, palloc(&p->lists, p->list_cap));
    P_HIP(c, hipHostMalloc((void **)&p->lists_h, sizeof(int32_t) * (size_t)p->list_cap, hipHostMallocDefault));
    P_HIP(c, palloc(&p->slot_ps, p->slot_cap));
    P_HIP(c, palloc(&p->slot_pm, p->slot_cap));
    P_HIP(c, palloc(&p->slot_lse, p->slot_cap));
    P_HIP(c, hipStreamSynchronize(st));
    p->rows_cap = (int)rows;
    return EMC_OK;
}

extern "C" emc_status emc_proposal_enable(emc_ctx *c, const float *coarse, int32_t n_coarse,
                                          const emc_proposal_options *opt)
{
    if (!c || !coarse || n_coarse < 1 || n_coarse > 8192)
        return pfail(c, EMC_E_INVALID, "proposal_enable: bad argument (1 <= n_coarse <= 8192)");
    emc_proposal_options o;
    if (opt) o = *opt;
    else emc_proposal_defaults(&o);
    if (!c->mfma || c->opt.diff != EMC_DIFF_TRUE_POISSON)
        return pfail(c, EMC_E_UNSUPPORTED, "proposal_enable: needs an MFMA precision (TRUE_POISSON)");
    if (c->opt.insert == EMC_INSERT_FUSED) return pfail(c, EMC_E_UNSUPPORTED, "proposal_enable: needs insert ATOMIC or BINNED");
    if (c->chunk < c->S) return pfail(c, EMC_E_UNSUPPORTED, "proposal_enable: slice_chunk must cover all rotations");
    if (!c->has_rot || !c->has_coords || !c->has_mask || !c->has_K)
        return pfail(c, EMC_E_STATE, "proposal_enable: needs rotations, coords, mask and patterns");
    if (c->k_split) return pfail(c, EMC_E_UNSUPPORTED, "proposal_enable: counts above 256 are not supported");
    if (c->chi2_mfma) return pfail(c, EMC_E_UNSUPPORTED, "proposal_enable: needs diff = TRUE_POISSON");
    if (o.top_k < 1 || o.top_k > 64 || o.top_k > n_coarse || o.hidden < 64 || o.hidden % 64 || o.hidden > 4096 ||
        o.batch < 64 || o.batch % 64 || o.batch > 65536 || o.epochs < 0 || o.warmup_full < 1 || o.full_every < 1 ||
        !(o.lr > 0.0f) || !(o.min_mass >= 0.0f) || !(o.max_fallback >= 0.0f) || o.synth_patterns < -1 ||
        o.refresh_epochs < 0 || o.refresh_epochs > 1000 || o.recall < 0 || o.recall > kMaxRecall ||
        (o.train_patterns != 0 && o.train_patterns != 1) || !(o.refresh_fraction > 0.0f && o.refresh_fraction <= 1.0f) ||
        !(o.train_below >= 0.0f && o.train_below <= 1.0f) ||
        o.synth_patterns > (1 << 22))
        return pfail(c, EMC_E_INVALID, "proposal_enable: bad option value");
    P_HIP(c, hipSetDevice(c->opt.device));
    P_TRY(emc_proposal_disable(c));
    c->prop = new Proposal();
    const emc_status s = enable_impl(c, coarse, n_coarse, o);
    if (s != EMC_OK) {
        const std::string msg = c->err;
        (void)emc_proposal_disable(c);
        c->err = msg;
    }
    return s;
}

extern "C" emc_status emc_proposal_predict(emc_ctx *c, int32_t *cells, float *mass)
{
    P_TRY(need_prop(c, "proposal_predict"));
    P_HIP(c, hipSetDevice(c->opt.device));
    P_TRY(proposal_predict(c));
    const Proposal *p = c->prop;
    if (cells) std::memcpy(cells, p->topc_h, sizeof(int32_t) * (size_t)c->I * p->o.top_k);
    if (mass) std::memcpy(mass, p->pmass_h, sizeof(float) * c->I);
    return EMC_OK;
}

extern "C" emc_status emc_proposal_captured(emc_ctx *c, double *cap)
{
    P_TRY(need_prop(c, "proposal_captured"));
    if (!cap) return pfail(c, EMC_E_INVALID, "proposal_captured: null argument");
    const Proposal *p = c->prop;
    if ((int)p->cap_h.size() != c->I) return pfail(c, EMC_E_STATE, "proposal_captured: no measurement yet");
    std::memcpy(cap, p->cap_h.data(), sizeof(double) * c->I);
    return EMC_OK;
}

extern "C" emc_status emc_proposal_recall(emc_ctx *c, int32_t *cells)
{
    P_TRY(need_prop(c, "proposal_recall"));
    if (!cells) return pfail(c, EMC_E_INVALID, "proposal_recall: null argument");
    const Proposal *p = c->prop;
    if (p->o.recall == 0) return EMC_OK;
    P_HIP(c, hipSetDevice(c->opt.device));
    P_HIP(c, hipMemcpyAsync(cells, p->recall, sizeof(int32_t) * (size_t)c->I * p->o.recall, hipMemcpyDeviceToHost,
                            c->stream));
    P_HIP(c, hipStreamSynchronize(c->stream));
    return EMC_OK;
}

extern "C" emc_status emc_proposal_cell_targets(emc_ctx *c, float *targets)
{
    P_TRY(need_prop(c, "proposal_cell_targets"));
    if (!targets) return pfail(c, EMC_E_INVALID, "proposal_cell_targets: null argument");
    if (!c->has_resp) return pfail(c, EMC_E_STATE, "proposal_cell_targets: needs full-search responsibilities");
    P_HIP(c, hipSetDevice(c->opt.device));
    P_TRY(proposal_targets(c, false, false, nullptr));
    const Proposal *p = c->prop;
    std::vector<float> Tt((size_t)p->C * c->Ip);
    P_HIP(c, hipMemcpyAsync(Tt.data(), p->Tt, sizeof(float) * Tt.size(), hipMemcpyDeviceToHost, c->stream));
    P_HIP(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->I; ++i)
        for (int k = 0; k < p->C; ++k) targets[(size_t)i * p->C + k] = Tt[(size_t)k * c->Ip + i];
    return EMC_OK;
}

extern "C" emc_status emc_proposal_train(emc_ctx *c, int32_t epochs, double *loss)
{
    P_TRY(need_prop(c, "proposal_train"));
    if (epochs < 0) return pfail(c, EMC_E_INVALID, "proposal_train: epochs must be >= 0");
    if (!c->has_resp) return pfail(c, EMC_E_STATE, "proposal_train: needs full-search responsibilities");
    P_HIP(c, hipSetDevice(c->opt.device));
    P_TRY(proposal_targets(c, false, false, nullptr));
    return proposal_train(c, epochs, true, 1.0, loss);
}

extern "C" emc_status emc_proposal_features(emc_ctx *c, float *raw)
{
    P_TRY(need_prop(c, "proposal_features"));
    if (!raw) return pfail(c, EMC_E_INVALID, "proposal_features: null argument");
    P_HIP(c, hipSetDevice(c->opt.device));
    Proposal *p = c->prop;
    P_HIP(c, mlp_features(nullptr, c->K8 ? nullptr : c->K16, c->K8, c->Pp, c->Pa, c->pix, c->L, c->I, c->Ip, p->raw,
                          c->stream));
    P_HIP(c, hipMemcpyAsync(raw, p->raw, sizeof(float) * (size_t)c->I * D, hipMemcpyDeviceToHost, c->stream));
    P_HIP(c, hipStreamSynchronize(c->stream));
    return EMC_OK;
}

extern "C" emc_status emc_mlp_get_params(emc_ctx *c, float *W1, float *b1, float *W2, float *b2, float *W3, float *b3)
{
    P_TRY(need_prop(c, "mlp_get_params"));
    P_HIP(c, hipSetDevice(c->opt.device));
    const Proposal *p = c->prop;
    std::vector<float> w(p->n_par);
    P_HIP(c, hipStreamSynchronize(c->stream));
    P_HIP(c, hipMemcpy(w.data(), p->par, sizeof(float) * w.size(), hipMemcpyDeviceToHost));
    const int H = p->H, C = p->C;
    if (W1) std::memcpy(W1, w.data() + p->oW1, sizeof(float) * (size_t)H * D);
    if (b1) std::memcpy(b1, w.data() + p->ob1, sizeof(float) * H);
    if (W2) std::memcpy(W2, w.data() + p->oW2, sizeof(float) * (size_t)H * H);
    if (b2) std::memcpy(b2, w.data() + p->ob2, sizeof(float) * H);
    if (W3) std::memcpy(W3, w.data() + p->oW3, sizeof(float) * (size_t)C * H);
    if (b3) std::memcpy(b3, w.data() + p->ob3, sizeof(float) * C);
    return EMC_OK;
}

extern "C" emc_status emc_mlp_set_params(emc_ctx *c, const float *W1, const float *b1, const float *W2, const float *b2,
                                         const float *W3, const float *b3)
{
    P_TRY(need_prop(c, "mlp_set_params"));
    if (!W1 || !b1 || !W2 || !b2 || !W3 || !b3) return pfail(c, EMC_E_INVALID, "mlp_set_params: null argument");
    P_HIP(c, hipSetDevice(c->opt.device));
    Proposal *p = c->prop;
    const int H = p->H, C = p->C;
    std::vector<float> w(p->n_par, 0.0f);
    std::memcpy(w.data() + p->oW1, W1, sizeof(float) * (size_t)H * D);
    std::memcpy(w.data() + p->ob1, b1, sizeof(float) * H);
    std::memcpy(w.data() + p->oW2, W2, sizeof(float) * (size_t)H * H);
    std::memcpy(w.data() + p->ob2, b2, sizeof(float) * H);
    std::memcpy(w.data() + p->oW3, W3, sizeof(float) * (size_t)C * H);
    std::memcpy(w.data() + p->ob3, b3, sizeof(float) * C);
    for (float v : w)
        if (!std::isfinite(v)) return pfail(c, EMC_E_INVALID, "mlp_set_params: non-finite parameter");
    P_HIP(c, hipStreamSynchronize(c->stream));
    P_HIP(c, hipMemcpy(p->par, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
    P_HIP(c, hipMemsetAsync(p->am, 0, sizeof(float) * p->n_par, c->stream));
    P_HIP(c, hipMemsetAsync(p->av, 0, sizeof(float) * p->n_par, c->stream));
    P_HIP(c, mlp_to_bf16(p->par, p->n_par, p->par16, c->stream));
    P_HIP(c, hipStreamSynchronize(c->stream));
    p->adam_t = 0;
    p->pred_valid = false;
    return EMC_OK;
}

extern "C" emc_status emc_mlp_forward(emc_ctx *c, const float *x, int32_t B, float *probs)
{
    P_TRY(need_prop(c, "mlp_forward"));
    if (!x || !probs) return pfail(c, EMC_E_INVALID, "mlp_forward: null argument");
    P_HIP(c, hipSetDevice(c->opt.device));
    Proposal *p = c->prop;
    P_TRY(upload_batch(c, x, nullptr, B));
    P_TRY(forward(c, p->Xb, p->o.batch, false));
    P_HIP(c, mlp_softmax_ce(p->z, p->Cp, p->C, B, p->o.batch, nullptr, 0.0f, p->prob, nullptr, nullptr, c->stream));
    std::vector<float> pr((size_t)p->o.batch * p->Cp);
    P_HIP(c, hipMemcpyAsync(pr.data(), p->prob, sizeof(float) * pr.size(), hipMemcpyDeviceToHost, c->stream));
    P_HIP(c, hipStreamSynchronize(c->stream));
    for (int b = 0; b < B; ++b)
        for (int k = 0; k < p->C; ++k) probs[(size_t)b * p->C + k] = pr[(size_t)b * p->Cp + k];
    return EMC_OK;
}

extern "C" emc_status emc_mlp_gradients(emc_ctx *c, const float *x, const float *t, int32_t B, double *loss, float *gW1,
                                        float *gb1, float *gW2, float *gb2, float *gW3, float *gb3)
{
    P_TRY(need_prop(c, "mlp_gradients"));
    if (!x || !t) return pfail(c, EMC_E_INVALID, "mlp_gradients: null argument");
    P_HIP(c, hipSetDevice(c->opt.device));
    Proposal *p = c->prop;
    P_TRY(upload_batch(c, x, t, B));
    P_HIP(c, launch_fill_f64(c->scal + SC_LOSS, 0.0, 1, c->stream));
    P_TRY(loss_and_grad(c, B));
    double l = 0.0;
    P_TRY(read_loss(c, &l));
    if (loss) *loss = l / B;
    std::vector<float> g(p->n_par);
    P_HIP(c, hipMemcpy(g.data(), p->grad, sizeof(float) * g.size(), hipMemcpyDeviceToHost));
    const int H = p->H, C = p->C;
    if (gW1) std::memcpy(gW1, g.data() + p->oW1, sizeof(float) * (size_t)H * D);
    if (gb1) std::memcpy(gb1, g.data() + p->ob1, sizeof(float) * H);
    if (gW2) std::memcpy(gW2, g.data() + p->oW2, sizeof(float) * (size_t)H * H);
    if (gb2) std::memcpy(gb2, g.data() + p->ob2, sizeof(float) * H);
    if (gW3) std::memcpy(gW3, g.data() + p->oW3, sizeof(float) * (size_t)C * H);
    if (gb3) std::memcpy(gb3, g.data() + p->ob3, sizeof(float) * C);
    return EMC_OK;
}

extern "C" emc_status emc_mlp_train_step(emc_ctx *c, const float *x, const float *t, int32_t B, double *loss)
{
    P_TRY(need_prop(c, "mlp_train_step"));
    if (!x || !t) return pfail(c, EMC_E_INVALID, "mlp_train_step: null argument");
    P_HIP(c, hipSetDevice(c->opt.device));
    P_TRY(upload_batch(c, x, t, B));
    P_HIP(c, launch_fill_f64(c->scal + SC_LOSS, 0.0, 1, c->stream));
    P_TRY(loss_and_grad(c, B));
    P_TRY(adam_step(c));
    double l = 0.0;
    P_TRY(read_loss(c, &l));
    if (loss) *loss = l / B;
    c->prop->pred_valid = false;
    return EMC_OK;
}
