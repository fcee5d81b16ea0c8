// Library context (opaque to callers of include/cipsi_ijkl.h).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <tuple>
#include <string>
#include <vector>

#include "common.h"

struct cipsi_ijkl_batch;

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct cipsi_ijkl_ctx {
    int device = 0;
    int num_cu = 256;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string last_error;
    // workspace (grow-only, device)
    DevBuf vsq;      // full symmetric AO pair matrix, later H^T
    DevBuf h;        // H[P_A][P_M]
    DevBuf cr;       // zero-padded C in 16-column blocks (cr_index)
    DevBuf packed;   // packed AO block built from labels
    DevBuf owner;    // last-writer election for duplicate labels
    DevBuf scal;     // small device scalars (first zero label, error flag)
    DevBuf scan;     // host entry points: input scan result (max |x| bits, non-finite, end)
    // host entry points: pinned bounce slots for the large transfers (allocated on first
    // use), each with the event of the last copy that used it
    static constexpr int kPinSlots = 4;
    static constexpr size_t kPinBytes = size_t(64) << 20;
    void* pin[kPinSlots] = {};
    hipEvent_t pin_ev[kPinSlots] = {};
    bool pin_busy[kPinSlots] = {};
    DevBuf io_in, io_in2, io_out, io_c;   // staging for host-pointer entry points
    // host entry points: the stream that copies finished parts of the MO stream to the
    // caller while half 2 still runs, and one event per half-2 segment (first use)
    static constexpr int kOutSegments = 16;
    hipStream_t copy_stream = nullptr;
    hipEvent_t seg_ev[kOutSegments] = {};
    // FP64 packed host entry point: the stream of the upload, the event of the last
    // expand run behind it, and the event of the context stream's earlier work
    hipStream_t up_stream = nullptr;
    hipEvent_t front_ev = nullptr;
    hipEvent_t entry_ev = nullptr;
    // Half-2 work tables, one per (N_MO, kind): kind > 0 the FP64 columns per work-group,
    // -1 the split path's triangle items, -2 the paired FP64 items of the whole half, -3
    // those of the host entry point's row ranges (seg: first item of each range).  Written once and kept
    // until release_workspace: a graph captured from an earlier call keeps its pointer.
    struct TileTable { DevBuf buf; int64_t total = 0; std::vector<int64_t> seg; };
    std::map<std::pair<int, int>, TileTable> tile_tables;
    // Buffers replaced by a larger allocation.  They are not freed before
    // release_workspace / destroy, so a captured graph that still names them stays valid.
    std::vector<DevBuf> retired;
    DevBuf slices;   // split-FP16 operand slices
    DevBuf adj_g;    // energy/gradient: packed MO integrals
    DevBuf adj_gh;   // energy/gradient: scaled pair density Gh[b][a] = m(b) Gamma(a,b)
    DevBuf adj_small;  // energy/gradient: h_MO, hC, reduction partials
    DevBuf cons;     // integral consumers: bad-label count, energy
    DevBuf sym_small;  // point-group path: stream end, basis tables, touched class quartets
    DevBuf sym_owner;  // point-group path: last-writer keys of the packed block
    DevBuf sym_ht;     // point-group path: H^T of every block
    DevBuf sym_tab;    // point-group path: per-block kernel arguments and item tables
    cipsi_ijkl_batch* batch_cache = nullptr;  // plan of the last cipsi_ijkl_transform_batch_dev
    // tile plans (kernel variant index) for the two halves: defaults and per-size plans
    int var_half1 = 2;
    int var_half2 = 2;
    std::map<std::tuple<int, int, int>, std::pair<int, int>> plans;
    // per-stage event timing
    bool timing = false;
    struct Pending { std::string stage; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::map<std::string, std::pair<double, int64_t>> totals;
};

// Event pairs held before they are folded into the per-stage totals (bounds the memory
// of a caller that enables timing and never reads it).
constexpr size_t kMaxPendingTimers = 4096;
void ctx_drain_timing(cipsi_ijkl_ctx* ctx);

// Records HIP events around one stage launch on the context stream when timing is on.
struct StageTimer {
    cipsi_ijkl_ctx* ctx;
    const char* stage;
    hipEvent_t a = nullptr;
    StageTimer(cipsi_ijkl_ctx* c, const char* s) : ctx(c), stage(s) {
        if (ctx->timing && hipEventCreate(&a) == hipSuccess) (void)hipEventRecord(a, ctx->stream);
        else a = nullptr;
    }
    ~StageTimer() {
        if (!a) return;
        hipEvent_t b;
        if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return; }
        (void)hipEventRecord(b, ctx->stream);
        ctx->pending.push_back({stage, a, b});
        if (ctx->pending.size() >= kMaxPendingTimers) ctx_drain_timing(ctx);
    }
};

int ctx_alloc(cipsi_ijkl_ctx* ctx, DevBuf& b, size_t bytes);
// Host -> device copy through the context's pinned bounce slots (host threads fill a
// slot while the DMA engine moves the previous one); ordered on the context stream.
int ctx_copy_h2d(cipsi_ijkl_ctx* ctx, void* d, const void* h, size_t bytes);
int ctx_enter(cipsi_ijkl_ctx* ctx);
int ctx_tile_table(cipsi_ijkl_ctx* ctx, int m, int lt, const int64_t** out, int64_t* total);
// A half-2 prefix table keyed by (m, kind): kind > 0 is the FP64 column width; the split
// path uses kind -1.  items(l, kind) is the number of work items of one MO pair of row l.
int ctx_prefix_table(cipsi_ijkl_ctx* ctx, std::pair<int, int> key, int m, int64_t (*items)(int, int),
                     const int64_t** out, int64_t* total);
int run_pad_c(cipsi_ijkl_ctx* ctx, int n, int m, const double* d_c, int ldc, double** cr);
// store_t (FP64 half 1): write H^T, rows of P_A per MO pair, instead of H.
int run_half(cipsi_ijkl_ctx* ctx, int half, int n, int m, const double* src, const double* cr,
             double* dst, int precision, int n_slices, int store_t = 0);
int run_half_variant(cipsi_ijkl_ctx* ctx, int half, int n, int m, const double* src,
                     const double* cr, double* dst, int precision, int n_slices, int var,
                     int store_t = 0);
int ctx_half_variant(const cipsi_ijkl_ctx* ctx, int half, int n, int m, int precision);

// No-allocation transform (noalloc.hip): resolves the device address of the static
// workspace on `device` (current) once per process; called by cipsi_ijkl_create.
int noalloc_setup(int device);

// energy / gradient (adjoint.hip)

// split-FP16 path (half_split.hip)
bool split_supported(int n_slices);
int64_t split_workspace_bytes(int n_ao, int n_mo);
double split_unit_roundoff(int n_ao, int n_slices);
int run_half_split(cipsi_ijkl_ctx* ctx, const cipsi::HalfArgs& a, int lt, int n_slices);
