// Whole FP64 transform of a packed AO block that allocates nothing.
//
// cipsi_ijkl_transform_packed_dev keeps its workspace (the symmetric pair matrix, H,
// the padded C and the half-2 work table) in grow-only buffers that the context
// allocates on first use.  This variant runs the same kernels, in the same order and
// with the same tile plan, on a workspace held in a file-scope device array instead, so
// a caller that may not allocate (code under graph capture, a serving loop, a harness
// that forbids hipMalloc) can call it at any time.  The array is sized for
// N_AO <= kNoAllocMaxAO; every call on one GPU shares that GPU's copy, so calls that
// overlap on one GPU (two streams, two host threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/cipsi_ijkl.h"
#include "common.h"
#include "ctx.h"

using namespace cipsi;

namespace {

constexpr int kNoAllocMaxAO = 48;
constexpr int64_t kNaPairs = (int64_t)kNoAllocMaxAO * (kNoAllocMaxAO + 1) / 2;   // 1176
constexpr int64_t kNaSq = kNaPairs * kNaPairs;           // doubles of the pair matrix / of H
constexpr int64_t kNaCr = 2 * (int64_t)kCRows * kNoAllocMaxAO;   // padded C and its second-product copy, N_MO <= N_AO
constexpr int64_t kNaTab = 64;                           // half-2 work table (N_MO + 1 <= 49)
// Layout in doubles: [pair matrix, later H^T | H | padded C, both orders | work table (int64)].
// Every region starts on a 256-byte boundary (kNaSq and kNaCr are multiples of 32).
constexpr int64_t kNaOffH = kNaSq;
constexpr int64_t kNaOffCr = 2 * kNaSq;
constexpr int64_t kNaOffTab = 2 * kNaSq + kNaCr;
constexpr int64_t kNaDoubles = kNaOffTab + kNaTab;       // 2,790,592 doubles = 22.3 MB
static_assert(kNaSq % 32 == 0 && kNaCr % 32 == 0, "regions must stay 256-byte aligned");

constexpr int kMaxDevices = 64;

}  // namespace

// The workspace.  In host code this name is only a handle for hipGetSymbolAddress: the
// device address of each GPU's copy is looked up once per device and kept in
// g_scratch_dev, which is what kernels and copies are given.
static __device__ __align__(256) double scratch[kNaDoubles];

static std::mutex g_scratch_mu;
static double* g_scratch_dev[kMaxDevices] = {};

// Half-2 work table tile_start[l] = sum_{l' < l} (l' + 1) ceil((l' + 1) / lt), written by
// the device so the entry point needs no host-to-device copy.
__global__ void k_noalloc_tile_table(int m, int lt, int64_t* __restrict__ t) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int64_t s = 0;
    t[0] = 0;
    for (int l = 0; l < m; ++l) {
        s += (int64_t)(l + 1) * ((l + lt) / lt);
        t[l + 1] = s;
    }
}

// Called by cipsi_ijkl_create with the context's device current: resolves that device's
// copy of the array once per process.
int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return CIPSI_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[device]) return CIPSI_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return CIPSI_ERR_HIP;
    }
    g_scratch_dev[device] = (double*)p;
    return CIPSI_OK;
}

static inline int round16(int x) { return (x + 15) & ~15; }

extern "C" int cipsi_ijkl_transform_packed_noalloc_dev(cipsi_ijkl_ctx* ctx, int n_ao, int n_mo,
                                                       const double* d_ao, const double* d_c,
                                                       int ldc, double* d_mo) {
    if (!ctx) return CIPSI_ERR_ARG;
    if (hipSetDevice(ctx->device) != hipSuccess) {
        (void)hipGetLastError();
        return CIPSI_ERR_HIP;
    }
    if (n_ao < 1 || n_mo < 1 || n_mo > n_ao || ldc < n_ao || !d_ao || !d_c || !d_mo)
        return CIPSI_ERR_ARG;
    if (n_ao > kNoAllocMaxAO) return CIPSI_ERR_UNSUPPORTED;
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return CIPSI_ERR_UNSUPPORTED;
    double* base = g_scratch_dev[ctx->device];       // set by cipsi_ijkl_create
    if (!base) return CIPSI_ERR_HIP;
    const int v1 = ctx_half_variant(ctx, 1, n_ao, n_mo, CIPSI_PREC_FP64);
    const int v2 = ctx_half_variant(ctx, 2, n_ao, n_mo, CIPSI_PREC_FP64);
    const int rc_cls = row_class(round16(n_ao));
    if (!variant_supports(v1, rc_cls) || !variant_supports(v2, rc_cls)) return CIPSI_ERR_UNSUPPORTED;

    const int64_t pa = npairs(n_ao), pm = npairs(n_mo);
    double* vsq = base;                  // pair matrix, then H^T
    double* h = base + kNaOffH;
    double* cr = base + kNaOffCr;
    int64_t* tab = (int64_t*)(base + kNaOffTab);
    hipStream_t st = ctx->stream;
    // every region is fully written before it is read, so the contents left by an
    // earlier call never matter: pad_c writes all kCRows x mp entries, expand the whole
    // pair matrix, half 1 every element of H, the transpose all of H^T
    hipError_t e = launch_pad_c(d_c, ldc, n_ao, n_mo, kCRows, round16(n_mo), cr, st);
    if (e == hipSuccess) e = launch_expand_packed(d_ao, pa, vsq, st);

    HalfArgs a{};
    a.src_ld = pa;
    a.cr = cr;
    a.ldm = round16(n_mo);
    a.n = n_ao;
    a.m = n_mo;
    a.np = round16(n_ao);
    a.mp = round16(n_mo);
    a.b_begin = 0;
    if (e == hipSuccess) {
        a.src = vsq;
        a.dst = h;
        a.half = 1;
        a.nb = pa;
        a.dst_ld = pm;
        const int lt = variant_columns(v1);
        a.ntile_total = pa * (int64_t)((n_mo + lt - 1) / lt);
        a.tile_start = nullptr;
        e = launch_half_fp64(a, v1, st);
    }
    if (e == hipSuccess) e = launch_transpose(h, pa, pm, vsq, st);
    if (e == hipSuccess) {
        const int lt = variant_columns(v2);
        hipLaunchKernelGGL(k_noalloc_tile_table, dim3(1), dim3(64), 0, st, n_mo, lt, tab);
        e = hipGetLastError();
        int64_t total = 0;
        for (int l = 0; l < n_mo; ++l) total += (int64_t)(l + 1) * ((l + lt) / lt);
        a.src = vsq;
        a.dst = d_mo;
        a.half = 2;
        a.nb = pm;
        a.dst_ld = 0;
        a.tile_start = tab;
        a.ntile_total = total;
        if (e == hipSuccess) e = launch_half_fp64(a, v2, st);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return e == hipErrorOutOfMemory ? CIPSI_ERR_NOMEM : CIPSI_ERR_HIP;
    }
    return CIPSI_OK;
}
