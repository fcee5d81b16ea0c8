// C ABI of the ijkl transform library (declarations: include/cipsi_ijkl.h).
//
// The pipeline for one C1 transform (all on one HIP stream, device-resident):
//   pad_c        C -> zero-padded copy in 16-column blocks used by the MFMA tiles
//   expand       packed AO block -> full symmetric pair matrix (rows contiguous)
//   half 1       per AO pair: C^T X C lower triangle   (tpqkld.f)      -> H^T (MO-pair
//                rows; replaces unit 40, itijkl.f:249-278)
//   half 2       per MO pair: C^T Y C, ij <= kl        (tijkld.f)      -> packed MO stream
#include <hip/hip_runtime.h>

#include <sys/mman.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cipsi_ijkl.h"
#include "common.h"
#include "ctx.h"

using namespace cipsi;

// --------------------------------------------------------------------------- helpers
static int map_hip(cipsi_ijkl_ctx* ctx, hipError_t e, const char* where) {
    if (e == hipSuccess) return CIPSI_OK;
    if (ctx) {
        ctx->last_error = std::string(where) + ": " + hipGetErrorString(e);
    }
    (void)hipGetLastError();
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return CIPSI_ERR_NOMEM;
    return CIPSI_ERR_HIP;
}

#define HIP_TRY(ctx, expr)                                          \
    do {                                                            \
        hipError_t _e = (expr);                                     \
        if (_e != hipSuccess) return map_hip((ctx), _e, #expr);     \
    } while (0)

// Grow-only workspace.  A buffer that is too small is retired, not freed: work already
// queued, or a graph captured from an earlier call, may still name it.  Retired buffers
// are freed by release_workspace / destroy.
int ctx_alloc(cipsi_ijkl_ctx* ctx, DevBuf& b, size_t bytes) {
    if (b.bytes >= bytes && b.ptr) return CIPSI_OK;
    if (bytes == 0) return CIPSI_OK;
    if (b.ptr) {
        try {
            ctx->retired.push_back(b);
        } catch (...) {
            return CIPSI_ERR_NOMEM;
        }
        b.ptr = nullptr;
        b.bytes = 0;
    }
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) { b.ptr = nullptr; return map_hip(ctx, e, "hipMalloc"); }
    b.bytes = bytes;
    return CIPSI_OK;
}

static void ctx_free(DevBuf& b) {
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

int ctx_enter(cipsi_ijkl_ctx* ctx) {
    if (!ctx) return CIPSI_ERR_ARG;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return map_hip(ctx, e, "hipSetDevice");
    return CIPSI_OK;
}

static int check_dims(int n_ao, int n_mo, int ldc) {
    if (n_ao < 1 || n_mo < 1 || n_mo > n_ao || ldc < n_ao) return CIPSI_ERR_ARG;
    if (n_ao > 256) return CIPSI_ERR_UNSUPPORTED;
    return CIPSI_OK;
}

static inline int round16(int x) { return (x + 15) & ~15; }

// Half-2 work tables: tile_start[l] = sum_{l'<l} (l'+1) * items(l'), items(l) the work
// items of one MO pair of row l.  One table per key, built on first use (a synchronous
// upload, so not under graph capture) and never rewritten, so a captured graph keeps a
// valid table whatever later calls use.
int ctx_prefix_table(cipsi_ijkl_ctx* ctx, std::pair<int, int> key, int m, int64_t (*items)(int, int),
                     const int64_t** out, int64_t* total) {
    auto it = ctx->tile_tables.find(key);
    if (it == ctx->tile_tables.end()) {
        std::vector<int64_t> t(m + 1);
        t[0] = 0;
        for (int l = 0; l < m; ++l) t[l + 1] = t[l] + (int64_t)(l + 1) * items(l, key.second);
        cipsi_ijkl_ctx::TileTable tab;
        int rc = ctx_alloc(ctx, tab.buf, sizeof(int64_t) * (m + 1));
        if (rc) return rc;
        hipError_t e = hipMemcpy(tab.buf.ptr, t.data(), sizeof(int64_t) * (m + 1),
                                 hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(tab.buf.ptr);
            return map_hip(ctx, e, "hipMemcpy(tile table)");
        }
        tab.total = t[m];
        try {
            it = ctx->tile_tables.emplace(key, tab).first;
        } catch (...) {
            (void)hipFree(tab.buf.ptr);
            return CIPSI_ERR_NOMEM;
        }
    }
    *out = (const int64_t*)it->second.buf.ptr;
    *total = it->second.total;
    return CIPSI_OK;
}

// FP64 half 2: an MO pair of row l has ceil((l+1)/lt) column tiles of lt columns.
static int64_t column_tiles(int l, int lt) { return (l + lt) / lt; }

int ctx_tile_table(cipsi_ijkl_ctx* ctx, int m, int lt, const int64_t** out, int64_t* total) {
    return ctx_prefix_table(ctx, std::make_pair(m, lt), m, column_tiles, out, total);
}

// Work items of the paired FP64 half 2 (build_pair_items) for N_MO = m, kept like the
// prefix tables: built and uploaded on first use, never rewritten.  kind -2 holds the whole
// half; kind -3 the row ranges [bound[s], bound[s + 1]) of the host entry point, items of
// range s from seg[s] (bound has nseg + 1 entries).
static int ctx_pair_table(cipsi_ijkl_ctx* ctx, int m, int kind, const int* bound, int nseg,
                          const PairItem** out, const cipsi_ijkl_ctx::TileTable** tab_out) {
    const auto key = std::make_pair(m, kind);
    auto it = ctx->tile_tables.find(key);
    if (it == ctx->tile_tables.end()) {
        std::vector<PairItem> items;
        std::vector<int64_t> seg;
        try {
            for (int r = 0; r < nseg; ++r) {
                seg.push_back((int64_t)items.size());
                build_pair_items(bound[r], bound[r + 1], items);
            }
            seg.push_back((int64_t)items.size());
        } catch (...) {
            return CIPSI_ERR_NOMEM;
        }
        cipsi_ijkl_ctx::TileTable tab;
        int rc = ctx_alloc(ctx, tab.buf, sizeof(PairItem) * std::max<size_t>(items.size(), 1));
        if (rc) return rc;
        hipError_t e = hipMemcpy(tab.buf.ptr, items.data(), sizeof(PairItem) * items.size(),
                                 hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(tab.buf.ptr);
            return map_hip(ctx, e, "hipMemcpy(pair table)");
        }
        tab.total = (int64_t)items.size();
        try {
            tab.seg = std::move(seg);
            it = ctx->tile_tables.emplace(key, std::move(tab)).first;
        } catch (...) {
            (void)hipFree(tab.buf.ptr);
            return CIPSI_ERR_NOMEM;
        }
    }
    *out = (const PairItem*)it->second.buf.ptr;
    *tab_out = &it->second;
    return CIPSI_OK;
}

int ctx_half_variant(const cipsi_ijkl_ctx* ctx, int half, int n, int m, int precision) {
    auto it = ctx->plans.find(std::make_tuple(n, m, precision));
    if (it != ctx->plans.end()) return half == 1 ? it->second.first : it->second.second;
    return half == 1 ? ctx->var_half1 : ctx->var_half2;
}

// --------------------------------------------------------------------------- pipeline pieces
int run_pad_c(cipsi_ijkl_ctx* ctx, int n, int m, const double* d_c, int ldc, double** cr) {
    const int np = round16(n), mp = round16(m);
    (void)np;
    int rc = ctx_alloc(ctx, ctx->cr, sizeof(double) * (size_t)cr_pair_size(mp));
    if (rc) return rc;
    StageTimer tm(ctx, "pad_c");
    HIP_TRY(ctx, launch_pad_c(d_c, ldc, n, m, kCRows, mp, (double*)ctx->cr.ptr, ctx->stream));
    *cr = (double*)ctx->cr.ptr;
    return CIPSI_OK;
}

int run_half(cipsi_ijkl_ctx* ctx, int half, int n, int m, const double* src, const double* cr,
             double* dst, int precision, int n_slices, int store_t) {
    return run_half_variant(ctx, half, n, m, src, cr, dst, precision, n_slices,
                            ctx_half_variant(ctx, half, n, m, precision), store_t);
}

int run_half_variant(cipsi_ijkl_ctx* ctx, int half, int n, int m, const double* src,
                     const double* cr, double* dst, int precision, int n_slices, int var,
                     int store_t) {
    if (precision == CIPSI_PREC_FP64 && !variant_supports(var, row_class(round16(n))))
        return CIPSI_ERR_UNSUPPORTED;
    if (store_t && half != 1) return CIPSI_ERR_ARG;
    const int64_t pa = npairs(n), pm = npairs(m);
    HalfArgs a{};
    a.src = src;
    a.src_ld = pa;
    a.cr = cr;
    a.ldm = round16(m);
    a.dst = dst;
    a.n = n;
    a.m = m;
    a.np = round16(n);
    a.mp = round16(m);
    a.half = half;
    a.b_begin = 0;
    a.store_t = store_t;
    const int lt = variant_columns(var);
    if (half == 1) {
        a.nb = pa;
        a.dst_ld = pm;
        a.ntile_total = pa * (int64_t)((m + lt - 1) / lt);
        a.tile_start = nullptr;
    } else {
        a.nb = pm;
        a.dst_ld = 0;
        const int64_t* ts;
        int64_t total;
        int rc = ctx_tile_table(ctx, m, lt, &ts, &total);
        if (rc) return rc;
        a.tile_start = ts;
        a.ntile_total = total;
    }
    const PairItem* pitems = nullptr;
    int64_t npitems = 0;
    const bool paired = precision == CIPSI_PREC_FP64 && half == 2 && var == kPairedHalf2Variant;
    if (paired) {
        const int whole[2] = {0, m};
        const cipsi_ijkl_ctx::TileTable* tab;
        int rc = ctx_pair_table(ctx, m, -2, whole, 1, &pitems, &tab);
        if (rc) return rc;
        npitems = tab->total;
    }
    StageTimer tm(ctx, half == 1 ? "half1" : "half2");
    if (paired) {
        HIP_TRY(ctx, launch_half2_pair(a, pitems, npitems, ctx->stream));
        return CIPSI_OK;
    }
    if (precision == CIPSI_PREC_FP64) {
        HIP_TRY(ctx, launch_half_fp64(a, var, ctx->stream));
        return CIPSI_OK;
    }
    if (precision == CIPSI_PREC_SPLIT_FP16) {
        return run_half_split(ctx, a, lt, n_slices);
    }
    return CIPSI_ERR_ARG;
}

static int check_prec(int precision, int n_slices) {
    if (precision == CIPSI_PREC_FP64) return CIPSI_OK;
    if (precision == CIPSI_PREC_SPLIT_FP16) return split_supported(n_slices) ? CIPSI_OK : CIPSI_ERR_UNSUPPORTED;
    return CIPSI_ERR_ARG;
}

// --------------------------------------------------------------------------- ABI
extern "C" {

const char* cipsi_ijkl_strerror(int s) {
    switch (s) {
        case CIPSI_OK: return "ok";
        case CIPSI_ERR_ARG: return "invalid argument (dimension, pointer or leading dimension)";
        case CIPSI_ERR_LABEL: return "label stream: N_AO > 255 or a label outside the AO block";
        case CIPSI_ERR_NOMEM: return "device memory allocation failed";
        case CIPSI_ERR_HIP: return "HIP runtime error";
        case CIPSI_ERR_NONFINITE: return "non-finite value in an input";
        case CIPSI_ERR_NOT_ORTHONORMAL: return "MO coefficients are not orthonormal";
        case CIPSI_ERR_UNSUPPORTED: return "unsupported size, precision or slice count";
        case CIPSI_ERR_NODEVICE: return "no usable gfx950 device";
        default: return "unknown status";
    }
}

const char* cipsi_ijkl_last_error(const cipsi_ijkl_ctx* ctx) {
    return ctx ? ctx->last_error.c_str() : "";
}

int cipsi_ijkl_create(cipsi_ijkl_ctx** out, int device) {
    if (!out) return CIPSI_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        (void)hipGetLastError();
        return CIPSI_ERR_NODEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return CIPSI_ERR_NODEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return CIPSI_ERR_NODEVICE;
    cipsi_ijkl_ctx* ctx = new (std::nothrow) cipsi_ijkl_ctx();
    if (!ctx) return CIPSI_ERR_NOMEM;
    ctx->device = device;
    ctx->num_cu = prop.multiProcessorCount;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return CIPSI_ERR_HIP;
    }
    ctx->stream = ctx->own_stream;
    if (int rc = noalloc_setup(device)) {       // the no-allocation entry point's workspace
        (void)hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return rc;
    }
    if (const char* v = std::getenv("CIPSI_IJKL_VARIANTS")) {     // "h1,h2" kernel variants
        int v1 = -1, v2 = -1;
        if (std::sscanf(v, "%d,%d", &v1, &v2) == 2 && v1 >= 0 && v1 < kNumHalfVariants && v2 >= 0 &&
            v2 < kNumHalfVariants) {
            ctx->var_half1 = v1;
            ctx->var_half2 = v2;
        }
    }
    *out = ctx;
    return CIPSI_OK;
}

int cipsi_ijkl_release_workspace(cipsi_ijkl_ctx* ctx) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->batch_cache) {
        cipsi_ijkl_batch_destroy(ctx, ctx->batch_cache);
        ctx->batch_cache = nullptr;
    }
    ctx_free(ctx->vsq);
    ctx_free(ctx->h);
    ctx_free(ctx->cr);
    ctx_free(ctx->packed);
    ctx_free(ctx->owner);
    ctx_free(ctx->io_in);
    ctx_free(ctx->io_in2);
    ctx_free(ctx->io_out);
    ctx_free(ctx->io_c);
    ctx_free(ctx->scal);
    ctx_free(ctx->scan);
    ctx_free(ctx->slices);
    ctx_free(ctx->adj_g);
    ctx_free(ctx->adj_gh);
    ctx_free(ctx->adj_small);
    ctx_free(ctx->cons);
    ctx_free(ctx->sym_small);
    ctx_free(ctx->sym_owner);
    ctx_free(ctx->sym_ht);
    ctx_free(ctx->sym_tab);
    for (auto& kv : ctx->tile_tables) ctx_free(kv.second.buf);
    ctx->tile_tables.clear();
    for (auto& b : ctx->retired) ctx_free(b);
    ctx->retired.clear();
    return CIPSI_OK;
}

static void pin_free(cipsi_ijkl_ctx* ctx);

int cipsi_ijkl_destroy(cipsi_ijkl_ctx* ctx) {
    if (!ctx) return CIPSI_ERR_ARG;
    cipsi_ijkl_reset_timing(ctx);
    cipsi_ijkl_release_workspace(ctx);
    pin_free(ctx);
    for (hipEvent_t& e : ctx->seg_ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    if (ctx->front_ev) (void)hipEventDestroy(ctx->front_ev);
    if (ctx->entry_ev) (void)hipEventDestroy(ctx->entry_ev);
    if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
    if (ctx->up_stream) (void)hipStreamDestroy(ctx->up_stream);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return CIPSI_OK;
}

int cipsi_ijkl_set_stream(cipsi_ijkl_ctx* ctx, void* s) {
    if (!ctx) return CIPSI_ERR_ARG;
    ctx->stream = (hipStream_t)s;
    return CIPSI_OK;
}

int cipsi_ijkl_reset_stream(cipsi_ijkl_ctx* ctx) {
    if (!ctx) return CIPSI_ERR_ARG;
    ctx->stream = ctx->own_stream;
    return CIPSI_OK;
}

int cipsi_ijkl_synchronize(cipsi_ijkl_ctx* ctx) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return CIPSI_OK;
}

}  // extern "C"

// A transform up to half 2: the padded C, the expanded pair matrix and half 1.  On return
// *ht is H^T (P_M rows of P_A) and *cr the padded C.  FP64 half 1 stores H^T itself: a
// lane's stores go to separate 128-byte lines, but each line holds 16 neighbouring AO
// pairs, which the chunked XCD mapping runs on one L2 at the same time, so the lines are
// completed there.  That costs half 1 about 2.3 ms at N = 240 and saves the 2.8 ms
// transpose.  The split path's last step of half 1 stores H^T the same way.
static int transform_front(cipsi_ijkl_ctx* ctx, int n_ao, int n_mo, const double* d_ao,
                           const double* d_c, int ldc, int precision, int n_slices,
                           const double** ht, double** cr) {
    const int64_t pa = npairs(n_ao), pm = npairs(n_mo);
    int rc;
    if ((rc = ctx_alloc(ctx, ctx->vsq, sizeof(double) * pa * pa))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->h, sizeof(double) * pa * pm))) return rc;
    if ((rc = run_pad_c(ctx, n_ao, n_mo, d_c, ldc, cr))) return rc;
    double* vsq = (double*)ctx->vsq.ptr;
    double* h = (double*)ctx->h.ptr;
    {
        StageTimer tm(ctx, "expand");
        HIP_TRY(ctx, launch_expand_packed(d_ao, pa, vsq, ctx->stream));
    }
    if ((rc = run_half(ctx, 1, n_ao, n_mo, vsq, *cr, h, precision, n_slices, 1))) return rc;
    *ht = h;
    return CIPSI_OK;
}

extern "C" {

int64_t cipsi_ijkl_workspace_bytes(int n_ao, int n_mo, int precision, int with_labels) {
    if (check_dims(n_ao, n_mo, n_ao)) return -1;
    const int64_t pa = npairs(n_ao), pm = npairs(n_mo);
    int64_t b = 8 * (pa * pa + pa * pm) + 8 * cr_pair_size(round16(n_mo));
    if (with_labels) b += 12 * npairs(pa);
    if (precision == CIPSI_PREC_SPLIT_FP16) b += split_workspace_bytes(n_ao, n_mo);
    return b;
}

int cipsi_ijkl_transform_packed_dev(cipsi_ijkl_ctx* ctx, int n_ao, int n_mo,
                                    const double* d_ao, const double* d_c, int ldc, double* d_mo,
                                    int precision, int n_slices) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_dims(n_ao, n_mo, ldc))) return rc;
    if ((rc = check_prec(precision, n_slices))) return rc;
    if (!d_ao || !d_c || !d_mo) return CIPSI_ERR_ARG;
    const double* ht;
    double* cr;
    if ((rc = transform_front(ctx, n_ao, n_mo, d_ao, d_c, ldc, precision, n_slices, &ht, &cr)))
        return rc;
    return run_half(ctx, 2, n_ao, n_mo, ht, cr, d_mo, precision, n_slices);
}

int cipsi_ijkl_unpack_labels_dev(cipsi_ijkl_ctx* ctx, int n_ao, const double* d_xx,
                                 const int32_t* d_labels, int64_t n_entries, double* d_ao) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if (n_ao < 1 || n_entries < 0 || !d_ao || (n_entries > 0 && (!d_xx || !d_labels)))
        return CIPSI_ERR_ARG;
    if (n_ao > 255) return CIPSI_ERR_LABEL;
    if (n_entries >= (int64_t)0xffffffffLL) return CIPSI_ERR_ARG;
    const int64_t nblk = npairs(npairs(n_ao));
    if ((rc = ctx_alloc(ctx, ctx->owner, sizeof(uint32_t) * nblk))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->scal, 64))) return rc;
    unsigned long long* sc = (unsigned long long*)ctx->scal.ptr;   // 4 scalars
    int* err_flag = (int*)((char*)ctx->scal.ptr + 32);
    {
        StageTimer tm(ctx, "unpack");
        HIP_TRY(ctx, launch_unpack_labels(d_xx, d_labels, n_entries, n_ao, d_ao,
                                          (uint32_t*)ctx->owner.ptr, sc, err_flag, ctx->stream));
    }
    int herr = 0;
    HIP_TRY(ctx, hipMemcpyAsync(&herr, err_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return herr ? CIPSI_ERR_LABEL : CIPSI_OK;
}

int cipsi_ijkl_transform_labels_dev(cipsi_ijkl_ctx* ctx, int n_ao, int n_mo, const double* d_xx,
                                    const int32_t* d_labels, int64_t n_entries, const double* d_c,
                                    int ldc, double* d_mo, int precision, int n_slices) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_dims(n_ao, n_mo, ldc))) return rc;
    if ((rc = check_prec(precision, n_slices))) return rc;
    const int64_t nblk = npairs(npairs(n_ao));
    if ((rc = ctx_alloc(ctx, ctx->packed, sizeof(double) * nblk))) return rc;
    if ((rc = cipsi_ijkl_unpack_labels_dev(ctx, n_ao, d_xx, d_labels, n_entries,
                                           (double*)ctx->packed.ptr)))
        return rc;
    return cipsi_ijkl_transform_packed_dev(ctx, n_ao, n_mo, (const double*)ctx->packed.ptr, d_c,
                                           ldc, d_mo, precision, n_slices);
}

int cipsi_ijkl_half1_dev(cipsi_ijkl_ctx* ctx, int n_ao, int n_mo, const double* d_ao,
                         const double* d_c, int ldc, double* d_h, int precision, int n_slices) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_dims(n_ao, n_mo, ldc))) return rc;
    if ((rc = check_prec(precision, n_slices))) return rc;
    if (!d_ao || !d_c || !d_h) return CIPSI_ERR_ARG;
    const int64_t pa = npairs(n_ao);
    if ((rc = ctx_alloc(ctx, ctx->vsq, sizeof(double) * pa * pa))) return rc;
    double* cr;
    if ((rc = run_pad_c(ctx, n_ao, n_mo, d_c, ldc, &cr))) return rc;
    {
        StageTimer tm(ctx, "expand");
        HIP_TRY(ctx, launch_expand_packed(d_ao, pa, (double*)ctx->vsq.ptr, ctx->stream));
    }
    return run_half(ctx, 1, n_ao, n_mo, (const double*)ctx->vsq.ptr, cr, d_h, precision, n_slices);
}

int cipsi_ijkl_half2_dev(cipsi_ijkl_ctx* ctx, int n_ao, int n_mo, const double* d_h,
                         const double* d_c, int ldc, double* d_mo, int precision, int n_slices) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_dims(n_ao, n_mo, ldc))) return rc;
    if ((rc = check_prec(precision, n_slices))) return rc;
    if (!d_h || !d_c || !d_mo) return CIPSI_ERR_ARG;
    const int64_t pa = npairs(n_ao), pm = npairs(n_mo);
    if ((rc = ctx_alloc(ctx, ctx->vsq, sizeof(double) * pa * pm))) return rc;
    double* cr;
    if ((rc = run_pad_c(ctx, n_ao, n_mo, d_c, ldc, &cr))) return rc;
    {
        StageTimer tm(ctx, "transpose");
        HIP_TRY(ctx, launch_transpose(d_h, pa, pm, (double*)ctx->vsq.ptr, ctx->stream));
    }
    return run_half(ctx, 2, n_ao, n_mo, (const double*)ctx->vsq.ptr, cr, d_mo, precision, n_slices);
}

// ------------------------------------------------------------------ host-pointer entry points
// ------------------------------------------------------------------ host transfers
// Large host <-> device copies of the host entry points go through four pinned 64 MiB
// slots: host threads copy a chunk between the caller's (pageable) memory and a slot
// while the DMA engine moves the previous slot, so the two overlap.  A slot is reused
// only after the event of its last DMA has completed.
// Up to 16 copy threads: the caller's buffer may be fresh memory, whose first touch
// (the kernel zeroing each page) halves a thread's copy rate, and the copy into it must
// keep up with the DMA engine.
static constexpr int kCopyThreads = 16;

static int copy_threads() {
    static const int n = [] {
        const unsigned hw = std::thread::hardware_concurrency();
        return hw == 0 ? 8 : (int)std::min<unsigned>(hw, kCopyThreads);
    }();
    return n;
}

static void parallel_memcpy(void* dst, const void* src, size_t bytes) {
    const int nthr = copy_threads();
    if (bytes < (size_t(4) << 20) || nthr < 2) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t part = ((bytes / nthr) + 63) & ~size_t(63);
    std::thread th[kCopyThreads];
    int used = 0;
    for (size_t off = 0; off < bytes; off += part) {
        const size_t len = std::min(part, bytes - off);
        try {           // no exception may leave the C ABI: copy here if no thread starts
            th[used] = std::thread([=] { std::memcpy((char*)dst + off, (const char*)src + off, len); });
            ++used;
        } catch (...) {
            std::memcpy((char*)dst + off, (const char*)src + off, len);
        }
    }
    for (int t = 0; t < used; ++t) th[t].join();
}

// Touches every page of the caller's output buffer from background threads.  Fresh
// memory faults on its first touch (the kernel zeroes each page), which halves the rate
// of the copy into it; touching it while the host waits for the upload and the
// transform takes that work off the copy.  join() must come before anything else writes
// the buffer; the destructor joins on every path.
static bool overlaps(const void* a, size_t na, const void* b, size_t nb) {
    const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
    return x < y + nb && y < x + na;
}

class PageToucher {
  public:
    // No touching when the output overlaps an input the call still has to read.
    PageToucher(void* p, size_t bytes, const void* in1, size_t n1, const void* in2 = nullptr,
                size_t n2 = 0, const void* in3 = nullptr, size_t n3 = 0) {
        if (overlaps(p, bytes, in1, n1) || (in2 && overlaps(p, bytes, in2, n2)) ||
            (in3 && overlaps(p, bytes, in3, n3)))
            return;
        // Ask for transparent huge pages over the whole 2 MiB pages inside the buffer: a
        // fresh 3.35 GB output then faults about 1,600 times instead of 820,000 (advice
        // only; without huge pages the touch below faults the small pages as before).
        constexpr uintptr_t kHuge = uintptr_t(2) << 20;
        const uintptr_t h0 = ((uintptr_t)p + kHuge - 1) & ~(kHuge - 1);
        const uintptr_t h1 = ((uintptr_t)p + bytes) & ~(kHuge - 1);
        if (h1 > h0) (void)madvise((void*)h0, h1 - h0, MADV_HUGEPAGE);
        constexpr size_t kPage = 4096;
        const size_t part = ((bytes / kThreads) + kPage - 1) & ~(kPage - 1);
        for (size_t off = 0; off < bytes && used_ < kThreads; off += part) {
            const size_t end = std::min(bytes, off + part);
            try {
                th_[used_] = std::thread([=] {
                    volatile char* q = (volatile char*)p;
                    for (size_t o = off; o < end; o += kPage) q[o] = 0;
                });
                ++used_;
            } catch (...) {
                break;                 // no thread: the copy faults the pages itself
            }
        }
    }
    void join() {
        for (int t = 0; t < used_; ++t) th_[t].join();
        used_ = 0;
    }
    ~PageToucher() { join(); }
    PageToucher(const PageToucher&) = delete;
    PageToucher& operator=(const PageToucher&) = delete;

  private:
    static constexpr int kThreads = 8;
    std::thread th_[kThreads];
    int used_ = 0;
};

static int pin_ready(cipsi_ijkl_ctx* ctx) {
    for (int s = 0; s < cipsi_ijkl_ctx::kPinSlots; ++s) {
        if (!ctx->pin[s]) {
            HIP_TRY(ctx, hipHostMalloc(&ctx->pin[s], cipsi_ijkl_ctx::kPinBytes, hipHostMallocDefault));
            HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->pin_ev[s], hipEventDisableTiming));
        }
    }
    return CIPSI_OK;
}

static int pin_wait(cipsi_ijkl_ctx* ctx, int s) {
    if (ctx->pin_busy[s]) {
        HIP_TRY(ctx, hipEventSynchronize(ctx->pin_ev[s]));
        ctx->pin_busy[s] = false;
    }
    return CIPSI_OK;
}

}  // extern "C"

int ctx_copy_h2d(cipsi_ijkl_ctx* ctx, void* d, const void* h, size_t bytes) {
    int rc = pin_ready(ctx);
    if (rc) return rc;
    const size_t chunk = cipsi_ijkl_ctx::kPinBytes;
    int s = 0;
    for (size_t off = 0; off < bytes; off += chunk, s = (s + 1) % cipsi_ijkl_ctx::kPinSlots) {
        const size_t len = std::min(chunk, bytes - off);
        if ((rc = pin_wait(ctx, s))) return rc;
        parallel_memcpy(ctx->pin[s], (const char*)h + off, len);
        HIP_TRY(ctx, hipMemcpyAsync((char*)d + off, ctx->pin[s], len, hipMemcpyHostToDevice,
                                    ctx->stream));
        HIP_TRY(ctx, hipEventRecord(ctx->pin_ev[s], ctx->stream));
        ctx->pin_busy[s] = true;
    }
    return CIPSI_OK;
}

extern "C" {

// Ordered after the work already queued on the context stream; returns when the caller's
// buffer holds the data.
static int copy_d2h(cipsi_ijkl_ctx* ctx, void* h, const void* d, size_t bytes,
                    PageToucher* touch = nullptr) {
    int rc = pin_ready(ctx);
    if (rc) return rc;
    constexpr int S = cipsi_ijkl_ctx::kPinSlots;
    const size_t chunk = cipsi_ijkl_ctx::kPinBytes;
    const int64_t nch = (int64_t)((bytes + chunk - 1) / chunk);
    auto issue = [&](int64_t k) -> int {
        const int s = (int)(k % S);
        const size_t off = (size_t)k * chunk, len = std::min(chunk, bytes - off);
        int r = pin_wait(ctx, s);
        if (r) return r;
        HIP_TRY(ctx, hipMemcpyAsync(ctx->pin[s], (const char*)d + off, len, hipMemcpyDeviceToHost,
                                    ctx->stream));
        HIP_TRY(ctx, hipEventRecord(ctx->pin_ev[s], ctx->stream));
        ctx->pin_busy[s] = true;
        return CIPSI_OK;
    };
    for (int64_t k = 0; k < nch && k < S; ++k)
        if ((rc = issue(k))) return rc;
    if (touch) touch->join();
    for (int64_t k = 0; k < nch; ++k) {
        const int s = (int)(k % S);
        const size_t off = (size_t)k * chunk, len = std::min(chunk, bytes - off);
        if ((rc = pin_wait(ctx, s))) return rc;
        parallel_memcpy((char*)h + off, ctx->pin[s], len);
        if (k + S < nch && (rc = issue(k + S))) return rc;
    }
    return CIPSI_OK;
}

// FP64 half 2 of a host entry point with the copy of its output to the caller running
// behind it.  The MO stream is ordered by MO pair row kl = pair0(l, k), so the rows with
// l < L are the prefix of npairs(npairs(L)) elements, and half 2's work items are ordered
// by l too (with the default tile, the paired items of each range of rows).  Half 2 is launched in up to kOutSegments ranges of l, each holding an equal
// share of the output and each followed by an event; a second stream copies the stream
// chunk by chunk through the pinned slots, every chunk after the event of the segment
// that completes it.  The PCIe copy (about 60 ms at N_MO = 240) then starts a few
// milliseconds into half 2 instead of after it.
static int half2_stream_out(cipsi_ijkl_ctx* ctx, int n, int m, const double* ht,
                            const double* cr, double* d_mo, double* mo, PageToucher* touch) {
    const int var = ctx_half_variant(ctx, 2, n, m, CIPSI_PREC_FP64);
    if (!variant_supports(var, row_class(round16(n)))) return CIPSI_ERR_UNSUPPORTED;
    const int lt = variant_columns(var);
    const int64_t* ts;
    int64_t total;
    int rc = ctx_tile_table(ctx, m, lt, &ts, &total);
    if (rc) return rc;
    if ((rc = pin_ready(ctx))) return rc;
    if (!ctx->copy_stream)
        HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
    for (hipEvent_t& e : ctx->seg_ev)
        if (!e) HIP_TRY(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // work items before row l (the device tile table, on the host)
    std::vector<int64_t> items(m + 1, 0);
    for (int l = 0; l < m; ++l) items[l + 1] = items[l] + (int64_t)(l + 1) * ((l + lt) / lt);
    auto out_before = [](int64_t l) { return npairs(npairs(l)); };   // elements of rows l' < l
    // segment s ends where the output reaches (s / K)^2 of the whole: the first segments
    // are small, so the copy starts early, and each later one adds a little more
    constexpr int K = cipsi_ijkl_ctx::kOutSegments;
    const int64_t nout = out_before(m);
    int bound[K + 1];
    int nseg = 0;
    bound[0] = 0;
    for (int s = 1; s <= K; ++s) {
        const int64_t goal = (int64_t)((double)nout * s * s / (K * K));
        int l = bound[nseg];
        while (l < m && out_before(l) < goal) ++l;
        if (l > bound[nseg]) bound[++nseg] = l;
    }
    const int64_t pa = npairs(n), pm = npairs(m);
    HalfArgs a{};
    a.src = ht;
    a.src_ld = pa;
    a.cr = cr;
    a.ldm = round16(m);
    a.dst = d_mo;
    a.n = n;
    a.m = m;
    a.np = round16(n);
    a.mp = round16(m);
    a.half = 2;
    a.nb = pm;
    a.tile_start = ts;
    // the default tile runs each range of rows with its column blocks paired (pair_tile)
    const PairItem* pitems = nullptr;
    const cipsi_ijkl_ctx::TileTable* ptab = nullptr;
    if (var == kPairedHalf2Variant && (rc = ctx_pair_table(ctx, m, -3, bound, nseg, &pitems, &ptab)))
        return rc;
    {
        StageTimer tm(ctx, "half2");
        for (int s = 0; s < nseg; ++s) {
            if (ptab) {
                HIP_TRY(ctx, launch_half2_pair(a, pitems + ptab->seg[s], ptab->seg[s + 1] - ptab->seg[s],
                                               ctx->stream));
            } else {
                a.item_begin = items[bound[s]];
                a.ntile_total = items[bound[s + 1]];
                HIP_TRY(ctx, launch_half_fp64(a, var, ctx->stream));
            }
            HIP_TRY(ctx, hipEventRecord(ctx->seg_ev[s], ctx->stream));
        }
    }
    // copy chunks: a slot each, doubling from 8 MiB over the first ones, so the copy starts
    // after the first segment
    constexpr int S = cipsi_ijkl_ctx::kPinSlots;
    const size_t bytes = sizeof(double) * (size_t)nout;
    std::vector<size_t> cut(1, 0);
    for (size_t len = size_t(8) << 20; cut.back() < bytes;
         len = std::min(cipsi_ijkl_ctx::kPinBytes, 2 * len))
        cut.push_back(std::min(bytes, cut.back() + len));
    const int64_t nch = (int64_t)cut.size() - 1;
    int waited = -1;                               // last segment the copy stream waits for
    auto issue = [&](int64_t k) -> int {
        const int s = (int)(k % S);
        const size_t off = cut[k], len = cut[k + 1] - cut[k];
        int r = pin_wait(ctx, s);
        if (r) return r;
        const int64_t end = (int64_t)((off + len) / sizeof(double));
        int g = waited < 0 ? 0 : waited;
        while (out_before(bound[g + 1]) < end) ++g;
        if (g > waited) {
            HIP_TRY(ctx, hipStreamWaitEvent(ctx->copy_stream, ctx->seg_ev[g], 0));
            waited = g;
        }
        HIP_TRY(ctx, hipMemcpyAsync(ctx->pin[s], (const char*)d_mo + off, len,
                                    hipMemcpyDeviceToHost, ctx->copy_stream));
        HIP_TRY(ctx, hipEventRecord(ctx->pin_ev[s], ctx->copy_stream));
        ctx->pin_busy[s] = true;
        return CIPSI_OK;
    };
    for (int64_t k = 0; k < nch && k < S; ++k)
        if ((rc = issue(k))) return rc;
    touch->join();
    for (int64_t k = 0; k < nch; ++k) {
        const int s = (int)(k % S);
        if ((rc = pin_wait(ctx, s))) return rc;
        parallel_memcpy((char*)mo + cut[k], ctx->pin[s], cut[k + 1] - cut[k]);
        if (k + S < nch && (rc = issue(k + S))) return rc;
    }
    return CIPSI_OK;
}

// The first packed row that starts at or after element e: the least a with npairs(a) >= e.
static int64_t first_row_from(int64_t e) {
    int64_t a = (int64_t)((std::sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
    while (a > 0 && npairs(a - 1) >= e) --a;
    while (npairs(a) < e) ++a;
    return a;
}

// The front of the FP64 packed host entry point: the upload of the packed AO block with
// the expand and half 1 run behind it.  Row b of the pair matrix is packed row b (columns
// s <= b) and column b of the packed rows after it (s > b), so it is complete once packed
// rows b .. P_A - 1 are on the device.  The block is therefore sent from its end, chunk by
// chunk through the pinned slots on the upload stream.  When a chunk completes the packed
// rows of more expand tile rows, those are expanded on the copy stream after the chunk's
// event, and half 1 runs on the context stream over the pair-matrix rows that every tile
// row from there to the end now covers.  Chunks shrink towards the start of the block,
// where one holds many short rows, so the last chunks complete few rows each and little
// of half 1 remains when the upload ends (at N = 240 the upload takes about 58 ms and
// half 1 about 25 ms).  The context stream's earlier work is ordered before the upload.
// On return every launch is queued; the upload stream holds the last copy.
static int upload_front(cipsi_ijkl_ctx* ctx, int n, int m, double* d_in, const double* ao,
                        double* vsq, const double* cr, double* ht) {
    const int var = ctx_half_variant(ctx, 1, n, m, CIPSI_PREC_FP64);
    if (!variant_supports(var, row_class(round16(n)))) return CIPSI_ERR_UNSUPPORTED;
    int rc = pin_ready(ctx);
    if (rc) return rc;
    if (!ctx->copy_stream)
        HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
    if (!ctx->up_stream)
        HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->up_stream, hipStreamNonBlocking));
    if (!ctx->front_ev) HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->front_ev, hipEventDisableTiming));
    if (!ctx->entry_ev) HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->entry_ev, hipEventDisableTiming));
    HIP_TRY(ctx, hipEventRecord(ctx->entry_ev, ctx->stream));
    HIP_TRY(ctx, hipStreamWaitEvent(ctx->up_stream, ctx->entry_ev, 0));
    const int64_t pa = npairs(n), pm = npairs(m);
    const int lt = variant_columns(var);
    const int64_t nt = (m + lt - 1) / lt;              // half-1 work items per matrix
    HalfArgs a{};
    a.src = vsq;
    a.src_ld = pa;
    a.cr = cr;
    a.ldm = round16(m);
    a.dst = ht;
    a.n = n;
    a.m = m;
    a.np = round16(n);
    a.mp = round16(m);
    a.half = 1;
    a.nb = pa;
    a.b_begin = 0;
    a.dst_ld = pm;
    a.store_t = 1;
    a.tile_start = nullptr;
    // chunk lengths in elements: about a quarter of what is left, between kMinChunk and a
    // slot, and doubling from kFirstChunk over the first chunks (the DMA engine starts
    // after a short host copy); a chunk starts on a 4 KiB boundary (rounded up, so it
    // never exceeds a slot)
    constexpr int64_t kSlot = (int64_t)(cipsi_ijkl_ctx::kPinBytes / sizeof(double));
    constexpr int64_t kMinChunk = int64_t(1) << 19;
    constexpr int64_t kFirstChunk = int64_t(1) << 20;
    constexpr int64_t kAlign = 512;
    // half 1 is launched once its next range holds this many matrices (or at the end):
    // about two rounds of work-groups over 256 CUs
    constexpr int64_t kMinRows = 256;
    const int64_t ntr = (pa + kExpandTile - 1) / kExpandTile;
    int64_t t_done = ntr;                              // tile rows [t_done, ntr) expanded
    int64_t b_done = pa;                               // half 1 queued for matrices >= b_done
    int64_t hi = npairs(pa);
    int64_t ramp = kFirstChunk;
    for (int s = 0; hi > 0; s = (s + 1) % cipsi_ijkl_ctx::kPinSlots) {
        const int64_t want = std::min({kSlot, ramp, std::max(kMinChunk, hi / 4)});
        ramp = std::min(kSlot, 2 * ramp);
        const int64_t lo = hi <= want ? 0 : (hi - want + kAlign - 1) / kAlign * kAlign;
        const size_t off = sizeof(double) * (size_t)lo, len = sizeof(double) * (size_t)(hi - lo);
        if ((rc = pin_wait(ctx, s))) return rc;
        parallel_memcpy(ctx->pin[s], (const char*)ao + off, len);
        HIP_TRY(ctx, hipMemcpyAsync((char*)d_in + off, ctx->pin[s], len, hipMemcpyHostToDevice,
                                    ctx->up_stream));
        HIP_TRY(ctx, hipEventRecord(ctx->pin_ev[s], ctx->up_stream));
        ctx->pin_busy[s] = true;
        hi = lo;
        // packed rows >= first_row_from(lo) have arrived: tile rows from t_ready on
        const int64_t t_ready = std::min(ntr, (first_row_from(lo) + kExpandTile - 1) / kExpandTile);
        if (t_ready < t_done) {
            HIP_TRY(ctx, hipStreamWaitEvent(ctx->copy_stream, ctx->pin_ev[s], 0));
            HIP_TRY(ctx, launch_expand_rows(d_in, pa, vsq, t_ready, t_done, ctx->copy_stream));
            t_done = t_ready;
            const int64_t b_ready = std::min(pa, t_ready * kExpandTile);
            if (b_done - b_ready >= kMinRows || b_ready == 0) {
                HIP_TRY(ctx, hipEventRecord(ctx->front_ev, ctx->copy_stream));
                HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream, ctx->front_ev, 0));
                a.item_begin = b_ready * nt;
                a.ntile_total = b_done * nt;
                HIP_TRY(ctx, launch_half_fp64(a, var, ctx->stream));
                b_done = b_ready;
            }
        }
    }
    return CIPSI_OK;                                   // the last chunk (lo = 0) queued the rest
}

// The device part of a host entry point from the resident packed block on: FP64 streams
// its output to the caller during half 2, the split path copies it after the transform.
static int transform_to_host(cipsi_ijkl_ctx* ctx, int n_ao, int n_mo, const double* d_ao,
                             const double* d_c, int ldc, int precision, int n_slices,
                             double* mo, PageToucher* touch) {
    double* d_mo = (double*)ctx->io_out.ptr;
    int rc;
    if (precision == CIPSI_PREC_FP64) {
        const double* ht;
        double* cr;
        if ((rc = transform_front(ctx, n_ao, n_mo, d_ao, d_c, ldc, precision, n_slices, &ht, &cr)))
            return rc;
        return half2_stream_out(ctx, n_ao, n_mo, ht, cr, d_mo, mo, touch);
    }
    if ((rc = cipsi_ijkl_transform_packed_dev(ctx, n_ao, n_mo, d_ao, d_c, ldc, d_mo, precision,
                                              n_slices)))
        return rc;
    return copy_d2h(ctx, mo, d_mo, sizeof(double) * npairs(npairs(n_mo)), touch);
}

static void pin_free(cipsi_ijkl_ctx* ctx) {
    for (int s = 0; s < cipsi_ijkl_ctx::kPinSlots; ++s) {
        if (ctx->pin_ev[s]) {
            (void)hipEventSynchronize(ctx->pin_ev[s]);
            (void)hipEventDestroy(ctx->pin_ev[s]);
            ctx->pin_ev[s] = nullptr;
        }
        if (ctx->pin[s]) (void)hipHostFree(ctx->pin[s]);
        ctx->pin[s] = nullptr;
        ctx->pin_busy[s] = false;
    }
}

static int check_finite_host(const double* p, int64_t n) {
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return CIPSI_ERR_NONFINITE;
    return CIPSI_OK;
}

static int upload_c(cipsi_ijkl_ctx* ctx, int n, int m, const double* c, int ldc, double** d_c) {
    int rc = ctx_alloc(ctx, ctx->io_c, sizeof(double) * (size_t)ldc * m);
    if (rc) return rc;
    for (int j = 0; j < m; ++j)
        if ((rc = check_finite_host(c + (int64_t)j * ldc, n))) return rc;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->io_c.ptr, c, sizeof(double) * (size_t)ldc * m,
                                hipMemcpyHostToDevice, ctx->stream));
    *d_c = (double*)ctx->io_c.ptr;
    return CIPSI_OK;
}

double err_estimate(int n, int precision, int n_slices, double vmax);

// Scan the device copy of the caller's values (finiteness, max |x|; up to the first zero
// label when labels are given) on stream st, after the upload, and wait for the result.
// A non-finite input is rejected before any output reaches the caller.  The scan runs at
// HBM speed instead of one host core's.
static int scan_uploaded(cipsi_ijkl_ctx* ctx, const double* d_p, int64_t n,
                         const int32_t* d_labels, double* vmax, hipStream_t st) {
    int rc = ctx_alloc(ctx, ctx->scan, 64);
    if (rc) return rc;
    unsigned long long* s = (unsigned long long*)ctx->scan.ptr;
    HIP_TRY(ctx, launch_scan_input(d_p, n, d_labels, s, st));
    unsigned long long hs[3] = {0ull, 0ull, 0ull};
    HIP_TRY(ctx, hipMemcpyAsync(hs, s, sizeof(hs), hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    if (hs[1]) return CIPSI_ERR_NONFINITE;
    double v;
    std::memcpy(&v, &hs[0], sizeof(v));
    *vmax = v;
    return CIPSI_OK;
}

int cipsi_ijkl_transform_packed(cipsi_ijkl_ctx* ctx, int n_ao, int n_mo, const double* ao,
                                const double* c, int ldc, double* mo, int precision, int n_slices,
                                double* err_est) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_dims(n_ao, n_mo, ldc))) return rc;
    if ((rc = check_prec(precision, n_slices))) return rc;
    if (!ao || !c || !mo) return CIPSI_ERR_ARG;
    const int64_t nin = npairs(npairs(n_ao)), nout = npairs(npairs(n_mo));
    if ((rc = ctx_alloc(ctx, ctx->io_in, sizeof(double) * nin))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->io_out, sizeof(double) * nout))) return rc;
    double* d_c;
    if ((rc = upload_c(ctx, n_ao, n_mo, c, ldc, &d_c))) return rc;
    PageToucher touch(mo, sizeof(double) * nout, ao, sizeof(double) * nin, c,
                      sizeof(double) * (size_t)ldc * n_mo);
    double vmax = 0.0;
    if (precision == CIPSI_PREC_FP64) {
        // the expand and half 1 run during the upload; the scan of the whole block still
        // rejects a non-finite one before half 2 and any copy to the caller
        const int64_t pa = npairs(n_ao), pm = npairs(n_mo);
        double* cr;
        if ((rc = ctx_alloc(ctx, ctx->vsq, sizeof(double) * pa * pa))) return rc;
        if ((rc = ctx_alloc(ctx, ctx->h, sizeof(double) * pa * pm))) return rc;
        if ((rc = run_pad_c(ctx, n_ao, n_mo, d_c, ldc, &cr))) return rc;
        if ((rc = upload_front(ctx, n_ao, n_mo, (double*)ctx->io_in.ptr, ao, (double*)ctx->vsq.ptr,
                               cr, (double*)ctx->h.ptr)))
            return rc;
        if ((rc = scan_uploaded(ctx, (const double*)ctx->io_in.ptr, nin, nullptr, &vmax,
                                ctx->up_stream))) {
            (void)hipStreamSynchronize(ctx->copy_stream);
            (void)hipStreamSynchronize(ctx->stream);
            return rc;
        }
        if ((rc = half2_stream_out(ctx, n_ao, n_mo, (const double*)ctx->h.ptr, cr,
                                   (double*)ctx->io_out.ptr, mo, &touch)))
            return rc;
    } else {
        if ((rc = ctx_copy_h2d(ctx, ctx->io_in.ptr, ao, sizeof(double) * nin))) return rc;
        if ((rc = scan_uploaded(ctx, (const double*)ctx->io_in.ptr, nin, nullptr, &vmax,
                                ctx->stream)))
            return rc;
        if ((rc = transform_to_host(ctx, n_ao, n_mo, (const double*)ctx->io_in.ptr, d_c, ldc,
                                    precision, n_slices, mo, &touch)))
            return rc;
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (err_est) *err_est = err_estimate(n_ao, precision, n_slices, vmax);
    return CIPSI_OK;
}

int cipsi_ijkl_transform_labels(cipsi_ijkl_ctx* ctx, int n_ao, int n_mo, const double* xx,
                                const int32_t* labels, int64_t n_entries, const double* c, int ldc,
                                double* mo, int precision, int n_slices, double* err_est) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_dims(n_ao, n_mo, ldc))) return rc;
    if (n_ao > 255) return CIPSI_ERR_LABEL;
    if ((rc = check_prec(precision, n_slices))) return rc;
    if (n_entries < 0 || !c || !mo || (n_entries > 0 && (!xx || !labels))) return CIPSI_ERR_ARG;
    const int64_t nout = npairs(npairs(n_mo));
    const int64_t ne = n_entries > 0 ? n_entries : 1;
    if ((rc = ctx_alloc(ctx, ctx->io_in, sizeof(double) * ne))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->io_in2, sizeof(int32_t) * ne))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->io_out, sizeof(double) * nout))) return rc;
    double* d_c;
    if ((rc = upload_c(ctx, n_ao, n_mo, c, ldc, &d_c))) return rc;
    PageToucher touch(mo, sizeof(double) * nout, xx, sizeof(double) * n_entries, labels,
                      sizeof(int32_t) * n_entries, c, sizeof(double) * (size_t)ldc * n_mo);
    if (n_entries > 0) {
        if ((rc = ctx_copy_h2d(ctx, ctx->io_in.ptr, xx, sizeof(double) * n_entries))) return rc;
        if ((rc = ctx_copy_h2d(ctx, ctx->io_in2.ptr, labels, sizeof(int32_t) * n_entries))) return rc;
    }
    double vmax = 0.0;          // 8 max |xx| before the first zero label (transd's largest scale)
    if ((rc = scan_uploaded(ctx, (const double*)ctx->io_in.ptr, n_entries,
                            (const int32_t*)ctx->io_in2.ptr, &vmax, ctx->stream)))
        return rc;
    vmax *= 8.0;
    if ((rc = ctx_alloc(ctx, ctx->packed, sizeof(double) * npairs(npairs(n_ao))))) return rc;
    if ((rc = cipsi_ijkl_unpack_labels_dev(ctx, n_ao, (const double*)ctx->io_in.ptr,
                                           (const int32_t*)ctx->io_in2.ptr, n_entries,
                                           (double*)ctx->packed.ptr)))
        return rc;
    if ((rc = transform_to_host(ctx, n_ao, n_mo, (const double*)ctx->packed.ptr, d_c, ldc,
                                precision, n_slices, mo, &touch)))
        return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (err_est) *err_est = err_estimate(n_ao, precision, n_slices, vmax);
    return CIPSI_OK;
}

int cipsi_ijkl_tmono(cipsi_ijkl_ctx* ctx, int n_ao, int n_mo, const double* h_packed,
                     const double* c, int ldc, double* h_mo) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_dims(n_ao, n_mo, ldc))) return rc;
    if (!h_packed || !c || !h_mo) return CIPSI_ERR_ARG;
    const int64_t pa = npairs(n_ao), pm = npairs(n_mo);
    if ((rc = check_finite_host(h_packed, pa))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->io_in, sizeof(double) * pa))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->io_out, sizeof(double) * pm))) return rc;
    double* d_c;
    if ((rc = upload_c(ctx, n_ao, n_mo, c, ldc, &d_c))) return rc;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->io_in.ptr, h_packed, sizeof(double) * pa,
                                hipMemcpyHostToDevice, ctx->stream));
    double* cr;
    if ((rc = run_pad_c(ctx, n_ao, n_mo, d_c, ldc, &cr))) return rc;
    HalfArgs a{};
    a.src = (const double*)ctx->io_in.ptr;
    a.src_ld = pa;
    a.cr = cr;
    a.ldm = round16(n_mo);
    a.dst = (double*)ctx->io_out.ptr;
    a.n = n_ao;
    a.m = n_mo;
    a.np = round16(n_ao);
    a.mp = round16(n_mo);
    a.half = 1;
    a.nb = 1;
    a.dst_ld = pm;
    a.ntile_total = (n_mo + variant_columns(0) - 1) / variant_columns(0);
    HIP_TRY(ctx, launch_half_fp64(a, 0, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(h_mo, ctx->io_out.ptr, sizeof(double) * pm, hipMemcpyDeviceToHost,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return CIPSI_OK;
}

extern "C++" void ctx_drain_timing(cipsi_ijkl_ctx* ctx) {
    for (auto& p : ctx->pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto& t = ctx->totals[p.stage];
            t.first += ms;
            t.second += 1;
        }
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    ctx->pending.clear();
}

int cipsi_ijkl_enable_timing(cipsi_ijkl_ctx* ctx, int enable) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    ctx->timing = enable != 0;
    return CIPSI_OK;
}

int cipsi_ijkl_reset_timing(cipsi_ijkl_ctx* ctx) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    ctx_drain_timing(ctx);
    ctx->totals.clear();
    return CIPSI_OK;
}

int cipsi_ijkl_read_timing(cipsi_ijkl_ctx* ctx, const char* stage, double* total_ms,
                           int64_t* launches) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if (!stage || !total_ms || !launches) return CIPSI_ERR_ARG;
    ctx_drain_timing(ctx);
    auto it = ctx->totals.find(stage);
    *total_ms = it == ctx->totals.end() ? 0.0 : it->second.first;
    *launches = it == ctx->totals.end() ? 0 : it->second.second;
    return CIPSI_OK;
}

}  // extern "C"

// A-priori error estimate: four chained length-N dot products with orthonormal C
// (column 2-norms 1), unit roundoff u of the arithmetic, random-walk growth sqrt(N).
double err_estimate(int n, int precision, int n_slices, double vmax) {
    double u = std::ldexp(1.0, -53);
    if (precision == CIPSI_PREC_SPLIT_FP16) u = split_unit_roundoff(n, n_slices);
    return 4.0 * std::sqrt((double)n) * u * vmax;
}
