// Tile autotuner for the half-transform kernels.
//
// For one (N_AO, N_MO, precision) it fills the workspace with synthetic data, times
// every built tile variant of half 1 and of half 2 with HIP events on the context
// stream (one warm-up launch, then the best of two timed launches) and records the
// fastest per half as the context's plan for that size.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "../../include/cipsi_ijkl.h"
#include "common.h"
#include "ctx.h"

using namespace cipsi;

namespace {

__global__ void k_fill_synthetic(double* __restrict__ p, int64_t n, uint64_t seed) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        uint64_t z = (uint64_t)i + seed * 0x9e3779b97f4a7c15ull;      // splitmix64 step
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        p[i] = (double)(z >> 11) * 0x1.0p-53 - 0.5;
    }
}

int fill(cipsi_ijkl_ctx* ctx, double* p, int64_t n, uint64_t seed) {
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_fill_synthetic, dim3((unsigned)g), dim3(256), 0, ctx->stream, p, n, seed);
    return hipGetLastError() == hipSuccess ? CIPSI_OK : CIPSI_ERR_HIP;
}

// Time one half with one variant; returns milliseconds (FLT_MAX on failure).
double time_variant(cipsi_ijkl_ctx* ctx, int half, int n, int m, const double* src,
                    const double* cr, double* dst, int precision, int var) {
    const int store_t = half == 1 && precision == CIPSI_PREC_FP64;   // as the transform runs
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return DBL_MAX;
    if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return DBL_MAX; }
    double best = DBL_MAX;
    if (run_half_variant(ctx, half, n, m, src, cr, dst, precision, 0, var, store_t) == CIPSI_OK) {
        for (int rep = 0; rep < 2; ++rep) {
            (void)hipEventRecord(a, ctx->stream);
            if (run_half_variant(ctx, half, n, m, src, cr, dst, precision, 0, var, store_t) !=
                CIPSI_OK)
                break;
            (void)hipEventRecord(b, ctx->stream);
            float ms = 0.f;
            if (hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess &&
                ms < best)
                best = ms;
        }
    }
    (void)hipGetLastError();
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return best;
}

}  // namespace

extern "C" {

int cipsi_ijkl_num_variants(void) { return kNumHalfVariants; }

int cipsi_ijkl_variant_info(int v, int* waves, int* kc, int* minb) {
    if (v < 0 || v >= kNumHalfVariants) return CIPSI_ERR_ARG;
    if (waves) *waves = kHalfVariants[v].waves;
    if (kc) *kc = kHalfVariants[v].kc;
    if (minb) *minb = kHalfVariants[v].minb;
    return CIPSI_OK;
}

int cipsi_ijkl_set_plan(cipsi_ijkl_ctx* ctx, int n_ao, int n_mo, int precision, int v1, int v2) {
    if (!ctx || v1 >= kNumHalfVariants || v2 >= kNumHalfVariants) return CIPSI_ERR_ARG;
    const int rc_cls = row_class((n_ao + 15) & ~15);
    if (precision == CIPSI_PREC_FP64 &&
        ((v1 >= 0 && !variant_supports(v1, rc_cls)) || (v2 >= 0 && !variant_supports(v2, rc_cls))))
        return CIPSI_ERR_UNSUPPORTED;
    auto key = std::make_tuple(n_ao, n_mo, precision);
    auto it = ctx->plans.find(key);
    std::pair<int, int> cur = it != ctx->plans.end() ? it->second
                                                     : std::make_pair(ctx->var_half1, ctx->var_half2);
    if (v1 >= 0) cur.first = v1;
    if (v2 >= 0) cur.second = v2;
    ctx->plans[key] = cur;
    return CIPSI_OK;
}

int cipsi_ijkl_get_plan(cipsi_ijkl_ctx* ctx, int n_ao, int n_mo, int precision, int* v1, int* v2) {
    if (!ctx) return CIPSI_ERR_ARG;
    if (v1) *v1 = ctx_half_variant(ctx, 1, n_ao, n_mo, precision);
    if (v2) *v2 = ctx_half_variant(ctx, 2, n_ao, n_mo, precision);
    return CIPSI_OK;
}

int cipsi_ijkl_autotune(cipsi_ijkl_ctx* ctx, int n_ao, int n_mo, int precision, int* best1,
                        int* best2) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if (n_ao < 1 || n_mo < 1 || n_mo > n_ao) return CIPSI_ERR_ARG;
    if (n_ao > 256) return CIPSI_ERR_UNSUPPORTED;
    if (precision != CIPSI_PREC_FP64) return CIPSI_ERR_UNSUPPORTED;
    const int64_t pa = npairs(n_ao), pm = npairs(n_mo);
    const int mp = (n_mo + 15) & ~15;
    if ((rc = ctx_alloc(ctx, ctx->vsq, sizeof(double) * pa * pa))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->h, sizeof(double) * pa * pm))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->io_out, sizeof(double) * npairs(pm)))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->cr, sizeof(double) * (size_t)cr_pair_size(mp)))) return rc;
    double* vsq = (double*)ctx->vsq.ptr;
    double* h = (double*)ctx->h.ptr;
    if ((rc = fill(ctx, vsq, pa * pa, 1))) return rc;
    if ((rc = fill(ctx, (double*)ctx->cr.ptr, cr_pair_size(mp), 2))) return rc;
    const double* cr = (const double*)ctx->cr.ptr;
    int v1 = ctx->var_half1, v2 = ctx->var_half2;
    double t1 = DBL_MAX, t2 = DBL_MAX;
    const int rcls = row_class((n_ao + 15) & ~15);
    for (int v = 0; v < kNumHalfVariants; ++v) {
        if (!variant_supports(v, rcls)) continue;
        double t = time_variant(ctx, 1, n_ao, n_mo, vsq, cr, h, precision, v);
        if (t < t1) { t1 = t; v1 = v; }
    }
    for (int v = 0; v < kNumHalfVariants; ++v) {
        if (!variant_supports(v, rcls)) continue;
        double t = time_variant(ctx, 2, n_ao, n_mo, vsq, cr, (double*)ctx->io_out.ptr, precision, v);
        if (t < t2) { t2 = t; v2 = v; }
    }
    if (t1 == DBL_MAX || t2 == DBL_MAX) return CIPSI_ERR_HIP;
    ctx->plans[std::make_tuple(n_ao, n_mo, precision)] = std::make_pair(v1, v2);
    if (best1) *best1 = v1;
    if (best2) *best2 = v2;
    return CIPSI_OK;
}

}  // extern "C"
