// Maximize, contraction half, sparse form: U[s, j] = sum_i P[s, i] K[i, j] over the patterns whose
// responsibility for rotation s is non-zero (emc_atomic.cu:88-132 computes the same sums densely).
//
// Once the model has formed, a pattern keeps its responsibility on a handful of rotations and
// exp() underflows to exactly 0 everywhere else (at 64^3 / 25,680 / 20,000 about 0.02 % of P is
// non-zero from the fifth iteration on), so the dense contraction multiplies almost only zeros.
// One workgroup per rotation row (sparse_row.h): the row of P is scanned, its non-zero entries are
// compacted in pattern order and their count rows summed.  Same operands as the dense path
// (P = hi + lo; counts as bytes, or BF16 [+ lo] above 255) with FP32 accumulation: only the order
// of the sum over patterns differs.
#include "../emc_internal.h"
#include "sparse_row.h"

namespace emc {
namespace {

template <int KT, int NPASS>
__global__ __launch_bounds__(256) void k_mstep_sparse(const bf16_t *__restrict__ Ph, const bf16_t *__restrict__ Pl,
                                                      int64_t ldp, int s0, int I, const void *__restrict__ Kv,
                                                      const bf16_t *__restrict__ Klo, int Pp,
                                                      const double *__restrict__ gate, double gate_thr,
                                                      float *__restrict__ U, const uint8_t *__restrict__ pdirty,
                                                      int nseg)
{
    if (gate && *gate > gate_thr) return;  // the dense contraction runs instead
    const size_t o = (size_t)(s0 + blockIdx.x) * ldp;
    sparse_row<KT, NPASS>(Ph + o, Pl + o, I, nullptr, Kv, Klo, Pp, U + (size_t)blockIdx.x * Pp,
                          pdirty ? pdirty + (size_t)(s0 + blockIdx.x) * nseg : nullptr);
}

}  // namespace

template <int NPASS>
static void launch_np(const emc_ctx *c, int s0, int cs, const double *gate, double gate_thr, hipStream_t st)
{
    const dim3 grid(cs), blk(256);
    if (c->K8)
        hipLaunchKernelGGL((k_mstep_sparse<0, NPASS>), grid, blk, 0, st, c->Ph, c->Pl, (int64_t)c->Ip, s0, c->I, c->K8,
                           nullptr, c->Pp, gate, gate_thr, c->U, c->pdirty, p_segments(c));
    else if (c->k_split)  // K16lo is kept allocated after a pattern set that needed it: key on k_split
        hipLaunchKernelGGL((k_mstep_sparse<2, NPASS>), grid, blk, 0, st, c->Ph, c->Pl, (int64_t)c->Ip, s0, c->I, c->K16,
                           c->K16lo, c->Pp, gate, gate_thr, c->U, c->pdirty, p_segments(c));
    else
        hipLaunchKernelGGL((k_mstep_sparse<1, NPASS>), grid, blk, 0, st, c->Ph, c->Pl, (int64_t)c->Ip, s0, c->I, c->K16,
                           nullptr, c->Pp, gate, gate_thr, c->U, c->pdirty, p_segments(c));
}

hipError_t launch_mstep_sparse(const emc_ctx *c, int s0, int cs, const double *gate, double gate_thr, hipStream_t st)
{
    if (c->Pp > kSparseMaxPp || c->Pp % 16 != 0) return hipErrorInvalidValue;
    if (c->Pp <= 4096) launch_np<1>(c, s0, cs, gate, gate_thr, st);
    else launch_np<kSparseMaxPp / 4096>(c, s0, cs, gate, gate_thr, st);
    return hipGetLastError();
}

}  // namespace emc
