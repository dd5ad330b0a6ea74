// CSC-segment prolongation fused with the id refactor.
//
// The reference expands every survivor's CSC column on the host into three
// per-output index arrays (24 B per output point), uploads them, runs an
// element-wise kernel and then a separate refactor kernel
// (SpMSpV.cu:200-254, SpMSpV.cuh:47-67, Refactor.cu:10-27).  Here one wave owns
// one survivor: its column is a contiguous run of nnz*D doubles in the P-tensor
// and in the output, so the wave streams it as a flat, fully coalesced array
// (y[e] = P[e] * x[e mod D]) and copies the matching child ids.  Survivor
// offsets come from the compaction's look-back, the survivor count from device
// memory.  HBM traffic per output point: 8D (P) + 8 (child id) in, 8D + 8 out.
#include "../ctx.hpp"

namespace tblaeq {

__global__ __launch_bounds__(256) void k_prolong(int D, const uint64_t* __restrict__ survIds,
                                                 const double* __restrict__ survVals,
                                                 const uint64_t* __restrict__ survOff,
                                                 const uint64_t* __restrict__ dSurvCount,
                                                 const uint64_t* __restrict__ colRes,
                                                 const double* __restrict__ pvals,
                                                 const uint64_t* __restrict__ childIds,
                                                 const double* __restrict__ packedP,
                                                 const uint64_t* __restrict__ packedChild,
                                                 uint64_t* __restrict__ outIds, double* __restrict__ outVals) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t wavesPerBlock = blockDim.x / kWave;
    const uint64_t gw = blockIdx.x * wavesPerBlock + threadIdx.x / kWave;
    const uint64_t nw = static_cast<uint64_t>(gridDim.x) * wavesPerBlock;
    const uint64_t ns = *dSurvCount;
    const int step = kWave % D;          // advance of (e mod D) per 64-element stride
    const int d0 = lane % D;
    for (uint64_t s = gw; s < ns; s += nw) {
        const uint64_t c = survIds[s];
        const uint64_t cs = colRes[c];
        const uint64_t cnt = colRes[c + 1] - cs;
        const uint64_t o = survOff[s];
        const uint64_t* __restrict__ src = packedChild ? packedChild + o : childIds + cs;
        uint64_t* __restrict__ dst = outIds + o;
        for (uint64_t j = lane; j < cnt; j += kWave) dst[j] = src[j];

        const double* __restrict__ P = packedP ? packedP + o * D : pvals + cs * D;
        const double* __restrict__ X = survVals + s * D;
        double* __restrict__ Y = outVals + o * D;
        const uint64_t total = cnt * static_cast<uint64_t>(D);
        if (step == 0) {
            const double xd = X[d0];
            uint64_t e = lane;
            for (; e + 3 * kWave < total; e += 4 * kWave) {
                const double p0 = P[e], p1 = P[e + kWave], p2 = P[e + 2 * kWave], p3 = P[e + 3 * kWave];
                Y[e] = p0 * xd;
                Y[e + kWave] = p1 * xd;
                Y[e + 2 * kWave] = p2 * xd;
                Y[e + 3 * kWave] = p3 * xd;
            }
            for (; e < total; e += kWave) Y[e] = P[e] * xd;
        } else {
            int d = d0;
            for (uint64_t e = lane; e < total; e += kWave) {
                Y[e] = P[e] * X[d];
                d += step;
                if (d >= D) d -= D;
            }
        }
    }
}

}  // namespace tblaeq

using namespace tblaeq;

extern "C" int tblaeq_prolong_refactor(tblaeq_ctx* ctx, const tblaeq_level* lv, const uint64_t* surv_ids,
                                       const double* surv_vals, const uint64_t* surv_off,
                                       const uint64_t* d_surv_count, uint64_t surv_bound, uint64_t* out_ids,
                                       double* out_vals) {
    if (!ctx || !lv || !surv_ids || !surv_vals || !surv_off || !d_surv_count || !out_ids || !out_vals ||
        lv->D == 0 || lv->D > 64) {
        set_error("tblaeq_prolong_refactor: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    const uint64_t waves = surv_bound == 0 ? 1 : surv_bound;
    uint64_t blocks = (waves + 3) / 4;
    const uint64_t cap = static_cast<uint64_t>(ctx->cuCount) * 16;
    if (blocks > cap) blocks = cap;
    TimedScope ts(ctx, kKProlong);
    hipLaunchKernelGGL(k_prolong, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, ctx->stream,
                       static_cast<int>(lv->D), surv_ids, surv_vals, surv_off, d_surv_count, lv->col_res, lv->pvals,
                       lv->child_ids, static_cast<const double*>(nullptr), static_cast<const uint64_t*>(nullptr),
                       out_ids, out_vals);
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}

extern "C" int tblaeq_prolong_refactor_packed(tblaeq_ctx* ctx, const tblaeq_level* lv, const uint64_t* surv_ids,
                                              const double* surv_vals, const uint64_t* surv_off,
                                              const uint64_t* d_surv_count, uint64_t surv_bound,
                                              const double* packed_pvals, const uint64_t* packed_child,
                                              uint64_t* out_ids, double* out_vals) {
    if (!ctx || !lv || !surv_ids || !surv_vals || !surv_off || !d_surv_count || !packed_pvals || !out_ids ||
        !out_vals || !lv->col_res || (!packed_child && !lv->child_ids) || lv->D == 0 || lv->D > 64) {
        set_error("tblaeq_prolong_refactor_packed: invalid argument");
        return TBLAEQ_ERR_INVALID;
    }
    const uint64_t waves = surv_bound == 0 ? 1 : surv_bound;
    uint64_t blocks = (waves + 3) / 4;
    const uint64_t cap = static_cast<uint64_t>(ctx->cuCount) * 16;
    if (blocks > cap) blocks = cap;
    TimedScope ts(ctx, kKProlong);
    hipLaunchKernelGGL(k_prolong, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, ctx->stream,
                       static_cast<int>(lv->D), surv_ids, surv_vals, surv_off, d_surv_count, lv->col_res,
                       static_cast<const double*>(nullptr), lv->child_ids, packed_pvals, packed_child, out_ids,
                       out_vals);
    TB_HIP(hipGetLastError());
    return TBLAEQ_OK;
}
