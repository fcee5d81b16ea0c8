// Batched transform of a scan of geometries with different basis sizes.
//
// The reference transforms one geometry per ijkl process, the scan driver starting
// one process per geometry (Autocip_13/autocip13.f90:525-567, ijkl.f:318).  Here the
// whole scan is one hipGraph launch:
//   pad_c (all)  ->  expand (all)  ->  half 1, one launch per row class, storing H^T
//   ->  half 2, one launch per row class (with the 8-wave, depth-16 variant: the paired
//       kernel of the single transform, its items numbered over the class)
// Every launch is grouped over the geometries: its kernels read the per-geometry
// arguments from device tables, so the captured graph does not depend on the caller's
// pointers and is reused by every run of the plan (a run uploads the tables, then
// launches the graph).  Geometries are grouped by row class (the number of 16-row
// blocks of N_AO, at least 4); the autotuner picks the tile variant of each half per
// class by timing the grouped launches on the plan's own data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/cipsi_ijkl.h"
#include "common.h"
#include "ctx.h"

using namespace cipsi;

namespace {

constexpr int kLts[4] = {32, 64, 128, 192};       // columns per work-group of the variants

int lt_index(int lt) { return lt == 32 ? 0 : (lt == 64 ? 1 : (lt == 128 ? 2 : 3)); }

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int round16(int x) { return (x + 15) & ~15; }

// Row stride of a geometry's expanded pair matrix: P_A rounded up to whole 128-byte lines,
// so that every row of the expand's stores and of half 1's Z panel loads starts on a line.
int64_t vsq_ld(int64_t pa) { return (pa + 15) & ~(int64_t)15; }

struct ClassPlan {
    int row_class = 0;
    std::vector<int> geoms;      // geometry indices, in scan order
    int first = 0;               // offset of this class in the per-class-ordered tables
    int v1 = 3, v2 = 3;          // tile variants of the two halves
    int64_t items1 = 0, items2 = 0;
    int64_t pair_first = 0;      // this class's paired half-2 items in pair_items
    int64_t npair = 0;
};

// Half 2 of a class runs paired (k_half2_pair_group) when its variant is the one the
// single transform runs paired.
bool paired_half2(const ClassPlan& cp) { return cp.v2 == kPairedHalf2Variant; }

}  // namespace

struct cipsi_ijkl_batch {
    int device = 0;
    int ngeom = 0;
    std::vector<int> n, m;
    std::vector<ClassPlan> classes;
    // device memory
    void* work = nullptr;        // vsq, h, cr of every geometry
    size_t work_bytes = 0;
    std::vector<size_t> off_vsq, off_h, off_cr;
    void* tables = nullptr;      // LayoutGeom[ngeom] | HalfArgs h1[ngeom] | HalfArgs h2[ngeom]
                                 // | start1[ngeom + ncls] | start2[ngeom + ncls]
    void* tile_tables = nullptr; // half-2 tile prefix tables, 4 column widths x sum(m + 1)
    std::vector<int64_t> tt_off; // offset of geometry g's table (per width: + w * tt_stride)
    int64_t tt_stride = 0;
    void* pair_items = nullptr;  // paired half-2 items of every class (GroupPairItem)
    size_t tables_bytes = 0;
    size_t off_layout = 0, off_h1 = 0, off_h2 = 0, off_s1 = 0, off_s2 = 0;
    void* host_tables = nullptr; // pinned staging of the tables
    hipEvent_t copied = nullptr; // last table upload done (staging reusable)
    bool copy_pending = false;
    int64_t padc_total = 0, expand_total = 0;
    // graph
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void drop_graph(cipsi_ijkl_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int fail(cipsi_ijkl_ctx* ctx, hipError_t e, const char* where) {
    ctx->last_error = std::string(where) + ": " + hipGetErrorString(e);
    (void)hipGetLastError();
    return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? CIPSI_ERR_NOMEM
                                                                       : CIPSI_ERR_HIP;
}

int check_sizes(int ngeom, const int* n_ao, const int* n_mo) {
    if (ngeom < 1 || !n_ao || !n_mo) return CIPSI_ERR_ARG;
    for (int g = 0; g < ngeom; ++g) {
        if (n_ao[g] < 1 || n_mo[g] < 1 || n_mo[g] > n_ao[g]) return CIPSI_ERR_ARG;
        if (n_ao[g] > 256) return CIPSI_ERR_UNSUPPORTED;
    }
    return CIPSI_OK;
}

// Fill the pinned host tables for these caller pointers and the current class plans.
void fill_tables(cipsi_ijkl_batch* b, const double* const* ao, const double* const* c,
                 const int* ldc, double* const* mo) {
    char* base = (char*)b->host_tables;
    LayoutGeom* lg = (LayoutGeom*)(base + b->off_layout);
    HalfArgs* h1 = (HalfArgs*)(base + b->off_h1);
    HalfArgs* h2 = (HalfArgs*)(base + b->off_h2);
    int64_t* s1 = (int64_t*)(base + b->off_s1);
    int64_t* s2 = (int64_t*)(base + b->off_s2);
    char* work = (char*)b->work;
    int64_t padc = 0, expd = 0;
    for (int g = 0; g < b->ngeom; ++g) {
        LayoutGeom& L = lg[g];
        L.ao = ao[g];
        L.c = c[g];
        L.vsq = (double*)(work + b->off_vsq[g]);
        L.h = (double*)(work + b->off_h[g]);
        L.cr = (double*)(work + b->off_cr[g]);
        L.n = b->n[g];
        L.m = b->m[g];
        L.np = round16(L.n);
        L.mp = round16(L.m);
        L.ldc = ldc[g];
        L.pa = npairs(L.n);
        L.pm = npairs(L.m);
        L.ldv = vsq_ld(L.pa);
        L.padc_start = padc;
        L.expand_start = expd;
        L.trans_start = 0;                       // half 1 stores H^T: no transpose
        padc += (int64_t)kCRows * L.mp;
        expd += expand_tiles(L.pa);
    }
    b->padc_total = padc;
    b->expand_total = expd;
    const int64_t* tt = (const int64_t*)b->tile_tables;
    int slot = 0;
    for (size_t k = 0; k < b->classes.size(); ++k) {
        ClassPlan& cp = b->classes[k];
        const int lt1 = variant_columns(cp.v1), lt2 = variant_columns(cp.v2);
        int64_t* st1 = s1 + cp.first + k;
        int64_t* st2 = s2 + cp.first + k;
        st1[0] = 0;
        st2[0] = 0;
        for (size_t i = 0; i < cp.geoms.size(); ++i, ++slot) {
            const int g = cp.geoms[i];
            const LayoutGeom& L = lg[g];
            HalfArgs a{};
            a.cr = L.cr;
            a.ldm = L.mp;
            a.n = L.n;
            a.m = L.m;
            a.np = L.np;
            a.mp = L.mp;
            a.src = L.vsq;
            a.src_ld = L.pa;
            a.b_begin = 0;
            HalfArgs x = a;                      // half 1: AO pairs of the expanded block,
            x.src_ld = L.ldv;                    // rows ldv apart, stored as H^T
            x.dst = L.h;
            x.store_t = 1;
            x.half = 1;
            x.nb = L.pa;
            x.dst_ld = L.pm;
            x.ntile_total = L.pa * (int64_t)((L.m + lt1 - 1) / lt1);
            x.tile_start = nullptr;
            h1[cp.first + i] = x;
            st1[i + 1] = st1[i] + x.ntile_total;
            HalfArgs y = a;                      // half 2: MO pairs of H^T
            y.src = L.h;
            y.dst = mo[g];
            y.half = 2;
            y.nb = L.pm;
            y.dst_ld = 0;
            const int64_t* tab = tt + lt_index(lt2) * b->tt_stride + b->tt_off[g];
            int64_t tot = 0;
            for (int l = 0; l < L.m; ++l) tot += (int64_t)(l + 1) * ((l + lt2) / lt2);
            y.ntile_total = tot;
            y.tile_start = tab;
            h2[cp.first + i] = y;
            st2[i + 1] = st2[i] + tot;
        }
        cp.items1 = st1[cp.geoms.size()];
        cp.items2 = paired_half2(cp) ? cp.npair : st2[cp.geoms.size()];
    }
}

int upload_tables(cipsi_ijkl_ctx* ctx, cipsi_ijkl_batch* b, const double* const* ao,
                  const double* const* c, const int* ldc, double* const* mo) {
    hipError_t e;
    if (b->copy_pending) {                       // staging still being read by a copy
        if ((e = hipEventSynchronize(b->copied)) != hipSuccess) return fail(ctx, e, "batch sync");
        b->copy_pending = false;
    }
    fill_tables(b, ao, c, ldc, mo);
    if ((e = hipMemcpyAsync(b->tables, b->host_tables, b->tables_bytes, hipMemcpyHostToDevice,
                            ctx->stream)) != hipSuccess)
        return fail(ctx, e, "batch table upload");
    if ((e = hipEventRecord(b->copied, ctx->stream)) != hipSuccess)
        return fail(ctx, e, "batch event");
    b->copy_pending = true;
    return CIPSI_OK;
}

GroupLaunch group_args(const cipsi_ijkl_batch* b, size_t k, int half) {
    const ClassPlan& cp = b->classes[k];
    const char* base = (const char*)b->tables;
    GroupLaunch gl{};
    gl.geoms = (const HalfArgs*)(base + (half == 1 ? b->off_h1 : b->off_h2)) + cp.first;
    gl.item_start = (const int64_t*)(base + (half == 1 ? b->off_s1 : b->off_s2)) + cp.first + k;
    gl.ngeom = (int)cp.geoms.size();
    gl.row_class = cp.row_class;
    gl.variant = half == 1 ? cp.v1 : cp.v2;
    if (half == 2 && paired_half2(cp)) {
        gl.paired = 1;
        gl.pair_items = (const GroupPairItem*)b->pair_items + cp.pair_first;
        gl.npair = cp.npair;
    }
    const int64_t items = half == 1 ? cp.items1 : cp.items2;
    gl.grid = xcd_grid(items);
    return gl;
}

int prepare_class(cipsi_ijkl_ctx* ctx, ClassPlan& cp) {
    GroupLaunch gl{};
    gl.row_class = cp.row_class;
    gl.grid = 0;
    gl.variant = cp.v1;
    hipError_t e = launch_half_fp64_group(gl, nullptr);
    if (e != hipSuccess) return fail(ctx, e, "batch prepare");
    gl.variant = cp.v2;
    gl.paired = paired_half2(cp);
    e = launch_half_fp64_group(gl, nullptr);
    if (e != hipSuccess) return fail(ctx, e, "batch prepare");
    return CIPSI_OK;
}

const LayoutGeom* layout_table(const cipsi_ijkl_batch* b) {
    return (const LayoutGeom*)((const char*)b->tables + b->off_layout);
}

// The launch sequence of one run (eager or under capture).
hipError_t enqueue_all(const cipsi_ijkl_batch* b, hipStream_t st) {
    const LayoutGeom* lg = layout_table(b);
    hipError_t e;
    if ((e = launch_group_pad_c(lg, b->ngeom, b->padc_total, st)) != hipSuccess) return e;
    if ((e = launch_group_expand(lg, b->ngeom, b->expand_total, st)) != hipSuccess) return e;
    for (size_t k = 0; k < b->classes.size(); ++k)
        if ((e = launch_half_fp64_group(group_args(b, k, 1), st)) != hipSuccess) return e;
    for (size_t k = 0; k < b->classes.size(); ++k)
        if ((e = launch_half_fp64_group(group_args(b, k, 2), st)) != hipSuccess) return e;
    return hipSuccess;
}

int capture(cipsi_ijkl_ctx* ctx, cipsi_ijkl_batch* b) {
    drop_graph(b);
    hipError_t e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return fail(ctx, e, "hipStreamBeginCapture");
    hipError_t le = enqueue_all(b, b->cap_stream);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->cap_stream, &g);
    if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(ctx, le, "batch capture launch");
    }
    if (e != hipSuccess) return fail(ctx, e, "hipStreamEndCapture");
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(b);
        return fail(ctx, e, "hipGraphInstantiate");
    }
    return CIPSI_OK;
}

int check_ptrs(const cipsi_ijkl_batch* b, const double* const* ao, const double* const* c,
               const int* ldc, double* const* mo) {
    if (!ao || !c || !ldc || !mo) return CIPSI_ERR_ARG;
    for (int g = 0; g < b->ngeom; ++g)
        if (!ao[g] || !c[g] || !mo[g] || ldc[g] < b->n[g]) return CIPSI_ERR_ARG;
    return CIPSI_OK;
}

// Time one grouped half launch (ms): one warm-up, then the best of two.
int time_group(cipsi_ijkl_ctx* ctx, const GroupLaunch& gl, float* ms) {
    hipEvent_t a, z;
    if (hipEventCreate(&a) != hipSuccess) return CIPSI_ERR_HIP;
    if (hipEventCreate(&z) != hipSuccess) { (void)hipEventDestroy(a); return CIPSI_ERR_HIP; }
    hipError_t e = launch_half_fp64_group(gl, ctx->stream);
    float best = 1e30f;
    for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
        (void)hipEventRecord(a, ctx->stream);
        e = launch_half_fp64_group(gl, ctx->stream);
        (void)hipEventRecord(z, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(z);
        float t = 0.f;
        if (e == hipSuccess && hipEventElapsedTime(&t, a, z) == hipSuccess && t < best) best = t;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(z);
    if (e != hipSuccess) return fail(ctx, e, "batch autotune");
    *ms = best;
    return CIPSI_OK;
}

}  // namespace

extern "C" {

int64_t cipsi_ijkl_batch_workspace_bytes(int n_geom, const int* n_ao, const int* n_mo) {
    if (check_sizes(n_geom, n_ao, n_mo)) return -1;
    int64_t t = 0;
    for (int g = 0; g < n_geom; ++g) {
        const int64_t pa = npairs(n_ao[g]), pm = npairs(n_mo[g]);
        t += align256(8 * pa * vsq_ld(pa)) + align256(8 * pa * pm) +
             align256(8 * (size_t)cr_pair_size(round16(n_mo[g])));
    }
    return t;
}

int cipsi_ijkl_batch_destroy(cipsi_ijkl_ctx* ctx, cipsi_ijkl_batch* b) {
    if (!b) return CIPSI_ERR_ARG;
    if (ctx) (void)hipSetDevice(ctx->device);
    if (b->copy_pending) (void)hipEventSynchronize(b->copied);
    if (b->cap_stream) (void)hipStreamSynchronize(b->cap_stream);
    drop_graph(b);
    if (b->work) (void)hipFree(b->work);
    if (b->tables) (void)hipFree(b->tables);
    if (b->tile_tables) (void)hipFree(b->tile_tables);
    if (b->pair_items) (void)hipFree(b->pair_items);
    if (b->host_tables) (void)hipHostFree(b->host_tables);
    if (b->copied) (void)hipEventDestroy(b->copied);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    delete b;
    return CIPSI_OK;
}

int cipsi_ijkl_batch_create(cipsi_ijkl_ctx* ctx, int n_geom, const int* n_ao, const int* n_mo,
                            cipsi_ijkl_batch** out) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if (!out) return CIPSI_ERR_ARG;
    *out = nullptr;
    if ((rc = check_sizes(n_geom, n_ao, n_mo))) return rc;
    cipsi_ijkl_batch* b = new (std::nothrow) cipsi_ijkl_batch();
    if (!b) return CIPSI_ERR_NOMEM;
    b->device = ctx->device;
    b->ngeom = n_geom;
    b->n.assign(n_ao, n_ao + n_geom);
    b->m.assign(n_mo, n_mo + n_geom);
    // classes, in increasing row class; geometries keep scan order inside a class
    int first = 0;
    for (int rc_ = 4; rc_ <= 16; ++rc_) {
        ClassPlan cp;
        cp.row_class = rc_;
        for (int g = 0; g < n_geom; ++g)
            if (row_class(round16(b->n[g])) == rc_) cp.geoms.push_back(g);
        if (cp.geoms.empty()) continue;
        cp.first = first;
        first += (int)cp.geoms.size();
        b->classes.push_back(cp);
    }
    const int ncls = (int)b->classes.size();
    // workspace offsets
    size_t off = 0;
    for (int g = 0; g < n_geom; ++g) {
        const int64_t pa = npairs(b->n[g]), pm = npairs(b->m[g]);
        b->off_vsq.push_back(off);
        off += align256(8 * pa * vsq_ld(pa));
        b->off_h.push_back(off);
        off += align256(8 * pa * pm);
        b->off_cr.push_back(off);
        off += align256(8 * (size_t)cr_pair_size(round16(b->m[g])));
    }
    b->work_bytes = off;
    // tables
    size_t t = 0;
    b->off_layout = t;
    t = align256(t + sizeof(LayoutGeom) * n_geom);
    b->off_h1 = t;
    t = align256(t + sizeof(HalfArgs) * n_geom);
    b->off_h2 = t;
    t = align256(t + sizeof(HalfArgs) * n_geom);
    b->off_s1 = t;
    t = align256(t + sizeof(int64_t) * (n_geom + ncls));
    b->off_s2 = t;
    t = align256(t + sizeof(int64_t) * (n_geom + ncls));
    b->tables_bytes = t;
    // half-2 tile prefix tables for the four column widths
    int64_t tts = 0;
    for (int g = 0; g < n_geom; ++g) {
        b->tt_off.push_back(tts);
        tts += b->m[g] + 1;
    }
    b->tt_stride = tts;
    std::vector<int64_t> tt(4 * tts);
    for (int w = 0; w < 4; ++w)
        for (int g = 0; g < n_geom; ++g) {
            int64_t* p = tt.data() + w * tts + b->tt_off[g];
            p[0] = 0;
            for (int l = 0; l < b->m[g]; ++l)
                p[l + 1] = p[l] + (int64_t)(l + 1) * ((l + kLts[w]) / kLts[w]);
        }
    // paired half-2 items, class after class (they depend on the sizes only)
    std::vector<GroupPairItem> pit;
    {
        std::vector<PairItem> one;
        for (auto& cp : b->classes) {
            cp.pair_first = (int64_t)pit.size();
            for (size_t i = 0; i < cp.geoms.size(); ++i) {
                one.clear();
                build_pair_items(0, b->m[cp.geoms[i]], one);
                for (const PairItem& it : one) pit.push_back(GroupPairItem{it, (int64_t)i});
            }
            cp.npair = (int64_t)pit.size() - cp.pair_first;
        }
    }
    hipError_t e;
    if ((e = hipMalloc(&b->work, b->work_bytes)) != hipSuccess ||
        (e = hipMalloc(&b->tables, b->tables_bytes)) != hipSuccess ||
        (e = hipMalloc(&b->tile_tables, sizeof(int64_t) * tt.size())) != hipSuccess ||
        (e = hipMalloc(&b->pair_items, sizeof(GroupPairItem) * pit.size())) != hipSuccess ||
        (e = hipHostMalloc(&b->host_tables, b->tables_bytes, hipHostMallocDefault)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->copied, hipEventDisableTiming)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking)) != hipSuccess) {
        rc = fail(ctx, e, "batch create");
        cipsi_ijkl_batch_destroy(ctx, b);
        return rc;
    }
    std::memset(b->host_tables, 0, b->tables_bytes);
    if ((e = hipMemcpy(b->tile_tables, tt.data(), sizeof(int64_t) * tt.size(),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->pair_items, pit.data(), sizeof(GroupPairItem) * pit.size(),
                       hipMemcpyHostToDevice)) != hipSuccess) {
        rc = fail(ctx, e, "batch create");
        cipsi_ijkl_batch_destroy(ctx, b);
        return rc;
    }
    for (auto& cp : b->classes) {
        cp.v1 = ctx->var_half1;
        cp.v2 = ctx->var_half2;
        if ((rc = prepare_class(ctx, cp))) {
            cipsi_ijkl_batch_destroy(ctx, b);
            return rc;
        }
    }
    *out = b;
    return CIPSI_OK;
}

int cipsi_ijkl_batch_run(cipsi_ijkl_ctx* ctx, cipsi_ijkl_batch* b, const double* const* d_ao,
                         const double* const* d_c, const int* ldc, double* const* d_mo,
                         int use_graph) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if (!b || b->device != ctx->device) return CIPSI_ERR_ARG;
    if ((rc = check_ptrs(b, d_ao, d_c, ldc, d_mo))) return rc;
    if ((rc = upload_tables(ctx, b, d_ao, d_c, ldc, d_mo))) return rc;
    StageTimer tm(ctx, "batch");
    hipError_t e;
    if (use_graph) {
        if (!b->exec && (rc = capture(ctx, b))) return rc;
        if ((e = hipGraphLaunch(b->exec, ctx->stream)) != hipSuccess)
            return fail(ctx, e, "hipGraphLaunch");
        return CIPSI_OK;
    }
    if ((e = enqueue_all(b, ctx->stream)) != hipSuccess) return fail(ctx, e, "batch launch");
    return CIPSI_OK;
}

int cipsi_ijkl_batch_autotune(cipsi_ijkl_ctx* ctx, cipsi_ijkl_batch* b, const double* const* d_ao,
                              const double* const* d_c, const int* ldc, double* const* d_mo) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if (!b || b->device != ctx->device) return CIPSI_ERR_ARG;
    if ((rc = check_ptrs(b, d_ao, d_c, ldc, d_mo))) return rc;
    drop_graph(b);
    // inputs of half 1
    if ((rc = upload_tables(ctx, b, d_ao, d_c, ldc, d_mo))) return rc;
    hipError_t e;
    const LayoutGeom* lg = layout_table(b);
    if ((e = launch_group_pad_c(lg, b->ngeom, b->padc_total, ctx->stream)) != hipSuccess ||
        (e = launch_group_expand(lg, b->ngeom, b->expand_total, ctx->stream)) != hipSuccess)
        return fail(ctx, e, "batch autotune");
    for (int half = 1; half <= 2; ++half) {
        for (size_t k = 0; k < b->classes.size(); ++k) {
            ClassPlan& cp = b->classes[k];
            int best_v = half == 1 ? cp.v1 : cp.v2;
            float best = 1e30f;
            for (int v = 0; v < kNumHalfVariants; ++v) {
                if (!variant_supports(v, cp.row_class)) continue;
                (half == 1 ? cp.v1 : cp.v2) = v;
                if ((rc = prepare_class(ctx, cp))) return rc;
                if ((rc = upload_tables(ctx, b, d_ao, d_c, ldc, d_mo))) return rc;
                float ms = 0.f;
                if ((rc = time_group(ctx, group_args(b, k, half), &ms))) return rc;
                if (ms < best) { best = ms; best_v = v; }
            }
            (half == 1 ? cp.v1 : cp.v2) = best_v;
            if ((rc = prepare_class(ctx, cp))) return rc;
        }
        if ((rc = upload_tables(ctx, b, d_ao, d_c, ldc, d_mo))) return rc;
        // after half 1 the last timed variant has written H^T, the same H^T as the chosen
        // one (the variants differ only in tiling), which half 2's timings read
    }
    // leave the outputs of a full run with the chosen plan
    if ((e = enqueue_all(b, ctx->stream)) != hipSuccess) return fail(ctx, e, "batch autotune");
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return fail(ctx, e, "batch autotune");
    return CIPSI_OK;
}

// One-call form: the context keeps the plan of its last size list.  A new size list
// builds a plan and tunes it on this call's data (the autotuner leaves the outputs of a
// full run); the same size list again replays the captured graph.
int cipsi_ijkl_transform_batch_dev(cipsi_ijkl_ctx* ctx, int n_geom, const int* n_ao,
                                   const int* n_mo, const double* const* d_ao,
                                   const double* const* d_c, const int* ldc,
                                   double* const* d_mo, int precision, int n_slices) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if (precision == CIPSI_PREC_SPLIT_FP16) return CIPSI_ERR_UNSUPPORTED;
    if (precision != CIPSI_PREC_FP64 || n_slices != 0) return CIPSI_ERR_ARG;
    if ((rc = check_sizes(n_geom, n_ao, n_mo))) return rc;
    cipsi_ijkl_batch* b = ctx->batch_cache;
    const bool same = b && b->ngeom == n_geom && std::equal(n_ao, n_ao + n_geom, b->n.begin()) &&
                      std::equal(n_mo, n_mo + n_geom, b->m.begin());
    if (same) return cipsi_ijkl_batch_run(ctx, b, d_ao, d_c, ldc, d_mo, 1);
    if (b) {
        cipsi_ijkl_batch_destroy(ctx, b);
        ctx->batch_cache = nullptr;
    }
    if ((rc = cipsi_ijkl_batch_create(ctx, n_geom, n_ao, n_mo, &b))) return rc;
    ctx->batch_cache = b;
    return cipsi_ijkl_batch_autotune(ctx, b, d_ao, d_c, ldc, d_mo);
}

int cipsi_ijkl_batch_num_classes(const cipsi_ijkl_batch* b) {
    return b ? (int)b->classes.size() : CIPSI_ERR_ARG;
}

int cipsi_ijkl_batch_class_info(const cipsi_ijkl_batch* b, int k, int* row_blocks, int* n_geom,
                                int* var_half1, int* var_half2) {
    if (!b || k < 0 || k >= (int)b->classes.size()) return CIPSI_ERR_ARG;
    const ClassPlan& cp = b->classes[k];
    if (row_blocks) *row_blocks = cp.row_class;
    if (n_geom) *n_geom = (int)cp.geoms.size();
    if (var_half1) *var_half1 = cp.v1;
    if (var_half2) *var_half2 = cp.v2;
    return CIPSI_OK;
}

int cipsi_ijkl_batch_set_class_plan(cipsi_ijkl_ctx* ctx, cipsi_ijkl_batch* b, int k,
                                    int var_half1, int var_half2) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if (!b || k < 0 || k >= (int)b->classes.size() || var_half1 < 0 ||
        var_half1 >= kNumHalfVariants || var_half2 < 0 || var_half2 >= kNumHalfVariants)
        return CIPSI_ERR_ARG;
    ClassPlan& cp = b->classes[k];
    if (!variant_supports(var_half1, cp.row_class) || !variant_supports(var_half2, cp.row_class))
        return CIPSI_ERR_UNSUPPORTED;
    cp.v1 = var_half1;
    cp.v2 = var_half2;
    drop_graph(b);
    return prepare_class(ctx, cp);
}

}  // extern "C"
