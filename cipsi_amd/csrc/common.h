// Shared definitions of the MI355X (gfx950) ijkl transform library.
//
// Index conventions (0-based; the reference is 1-based):
//   pair0(a, b) = a(a+1)/2 + b, a >= b         (CIPSI_sources/ijkl/itijkl.f:38)
//   P_A = N(N+1)/2 AO pairs, P_M = M(M+1)/2 MO pairs
//   packed AO block   V[pair0(ij, kl)], ij >= kl            (transd.f:369-371)
//   half-transformed  H[pq * P_M + pair0(l, k)], l >= k      (tpqkld.f:91-102)
//   packed MO stream  out[pair0(kl, pair0(j, i))], ij <= kl   (tijkld.f:31-79)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace cipsi {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

__host__ __device__ inline int64_t pair0(int64_t a, int64_t b) { return a * (a + 1) / 2 + b; }
__host__ __device__ inline int64_t pairsym(int64_t a, int64_t b) { return a >= b ? pair0(a, b) : pair0(b, a); }
__host__ __device__ inline int64_t npairs(int64_t n) { return n * (n + 1) / 2; }

// Largest a with pair0(a, 0) <= x (row of packed index x).
__device__ inline int64_t pair_row(int64_t x) {
    int64_t a = (int64_t)((sqrt(8.0 * (double)x + 1.0) - 1.0) * 0.5);
    while (pair0(a, 0) > x) --a;
    while (pair0(a + 1, 0) <= x) ++a;
    return a;
}

// One half of the transform, as a batch of "sandwiches": for every matrix index b
//   Z_b(r, s) = src[b * src_ld + pair0(max(r,s), min(r,s))]  (symmetric n x n)
//   O_b = C^T Z_b C  restricted to a <= c and pair0(c, a) < limit_b, written to
//   dst[dst_off_b + pair0(c, a)].
// Half 1 (tpqkld): b = AO pair pq, src = full symmetric AO supermatrix rows,
//   limit = P_M, dst_off = pq * P_M.
// Half 2 (tijkld): b = MO pair kl, src = H^T rows, limit = kl + 1,
//   dst_off = pair0(kl, 0); only columns c <= l (kl = pair0(l, k)) are needed.
// Padded copy of C in 16-column blocks: block a / 16 holds all kCRows rows of its 16
// columns, row-major, so C(r, a) = cr[((a / 16) * kCRows + r) * 16 + a % 16].  Every
// row class (up to 16 blocks of 16 rows) reads rows < kCRows; rows >= N_AO and
// columns >= N_MO are zero.
constexpr int kCRows = 256;
__host__ __device__ inline int64_t cr_index(int r, int a) {
    return ((int64_t)(a >> 4) * kCRows + r) * 16 + (a & 15);
}
// Every buffer that holds the padded C holds a second copy right after it, kCRows * mp
// doubles further on (pad_c writes both; cr_pair_size gives the doubles of both).  The
// copy keeps the 16 x 16 blocks (row block rb, column block ab) but orders each one for
// the B operand of the f64 MFMA in the second product, where lane l = 16 lg + lm needs
// C(16 rb + 4 q + lg, 16 ab + lm) for q = 0..3: q = 2h + j sits at h * 128 + 2 l + j, so a
// lane loads its four values as two 16-byte pieces and a wave's load is 1 KiB contiguous.
__host__ __device__ inline int64_t cq_index(int r, int a) {
    const int rr = r & 15, q = rr >> 2, l = 16 * (rr & 3) + (a & 15);
    return ((int64_t)(a >> 4) * 16 + (r >> 4)) * 256 + (q >> 1) * 128 + 2 * l + (q & 1);
}
__host__ __device__ inline int64_t cr_pair_size(int mp) { return 2 * (int64_t)kCRows * mp; }

struct HalfArgs {
    const double* src;     // rows of packed symmetric matrices
    int64_t src_ld;        // = P_A
    const double* cr;      // C in 16-column blocks, zero padded (cr_index)
    int ldm;               // = mp (columns of the padded copy)
    double* dst;
    int n;                 // N_AO
    int m;                 // N_MO
    int np;                // N rounded up to 16
    int mp;                // M rounded up to 16
    int half;              // 1 or 2
    int64_t nb;            // number of matrices (P_A or P_M)
    int64_t b_begin;       // first matrix index handled by this launch
    int64_t dst_ld;        // half 1: P_M
    int64_t ntile_total;   // end of the (b, column tile) work items of this launch
    int64_t item_begin;    // first work item of this launch (0: the whole half)
    int store_t;           // half 1: 1 = write H^T (dst[pair0(c, a) * nb + b]) instead of
                           // H (dst[b * dst_ld + pair0(c, a)])
    const int64_t* tile_start;  // half 2: prefix table over l (m + 1 entries), device
};

// A tile variant of the half-transform kernel: waves per work-group (16 output
// columns each), Q1 panel depth, and the work-groups-per-CU occupancy hint.
struct HalfVariant {
    int waves;
    int kc;
    int minb;
};
extern const HalfVariant kHalfVariants[];
extern const int kNumHalfVariants;
inline int variant_columns(int v) { return 16 * kHalfVariants[v].waves; }
// Whether variant v is built for a row class (the number of 16-row blocks of N_AO).
bool variant_supports(int v, int row_cls);

// Device launchers (defined in the .hip files).
hipError_t launch_unpack_labels(const double* xx, const int32_t* labels, int64_t n_entries,
                                int n_ao, double* ao_packed, uint32_t* owner,
                                unsigned long long* sc, int* err_flag, hipStream_t st);
hipError_t launch_expand_packed(const double* ao_packed, int64_t pa, double* vsq, hipStream_t st);
// Rows of 64 x 64 expand tiles: tile row i covers pair-matrix rows 64 i .. 64 i + 63 and
// reads only packed rows 64 i .. 64 i + 63.
constexpr int kExpandTile = 64;
hipError_t launch_expand_rows(const double* ao_packed, int64_t pa, double* vsq, int64_t i0,
                              int64_t i1, hipStream_t st);
// Input scan for the host entry points: s[0] = bits of max |x|, s[1] = any non-finite,
// over p[0 .. end) with end = the first zero label's index (labels given) or n.  p must
// be 16-byte aligned; s holds three 64-bit words.
hipError_t launch_scan_input(const double* p, int64_t n, const int32_t* labels,
                             unsigned long long* s, hipStream_t st);
hipError_t launch_transpose(const double* src, int64_t rows, int64_t cols, double* dst,
                            hipStream_t st);
hipError_t launch_pad_c(const double* c, int ldc, int n, int m, int np, int mp, double* cr,
                        hipStream_t st);
hipError_t launch_half_fp64(const HalfArgs& a, int variant, hipStream_t st);

// Half 2 of one transform with the column blocks of its matrices regrouped into fuller work
// groups (half_fp64.hip): an item holds one or two pieces of matrices; slots byte w says what
// wave w does (bit 7 busy, bit 4 the piece, bits 0-3 the column block).
struct PairItem {
    int64_t b[2];          // matrices (MO pairs) of the pieces; b[1] < 0: one piece
    uint64_t slots;
};
hipError_t launch_half2_pair(const HalfArgs& a, const PairItem* items, int64_t nitems,
                             hipStream_t st);
void build_pair_items(int l0, int l1, std::vector<PairItem>& out);
// The tile variant (8 waves, panel depth 16) whose single-transform half 2 runs paired.
constexpr int kPairedHalf2Variant = 2;

// Row class (compile-time row-block count) used for a padded AO dimension np.
int row_class(int np);

// A paired half-2 item of a grouped launch: its pieces belong to transform geom.
struct GroupPairItem {
    PairItem p;
    int64_t geom;          // index into GroupLaunch::geoms
};

// A grouped launch of one half over several transforms whose padded AO dimension
// falls in one row class.  geoms[g] are the per-transform arguments (device
// array); item_start (device, ngeom + 1 entries) numbers their work items.
// paired != 0 (half 2, variant kPairedHalf2Variant): the work items are pair_items
// (device, npair entries), built by build_pair_items per transform.
struct GroupLaunch {
    const HalfArgs* geoms;
    const int64_t* item_start;
    int ngeom;
    int row_class;
    int variant;
    int64_t grid;    // > 0: launch with this grid (>= item total); <= 0: prepare only
    int paired;
    const GroupPairItem* pair_items;
    int64_t npair;
};
hipError_t launch_half_fp64_group(const GroupLaunch& gl, hipStream_t st);
// XCD-aware work-item mapping of the half-transform kernels.  The hardware deals
// work-groups to the 8 XCDs round robin (work-group i runs on XCD i % 8).  Work items
// are cut into chunks of kXcdChunk consecutive items and the chunks are dealt to the
// XCDs round robin too: the tiles of one matrix (consecutive items) share an L2, and
// every XCD gets the same mix of cheap and expensive items (in half 2 an item's cost
// grows with its MO pair row; with contiguous eighths a cost model of the items gives
// the last XCD 10-17 % more work than the mean).  xcd_grid rounds a launch up to a multiple of 8 kXcdChunk work-groups,
// which makes xcd_item a permutation of [0, grid).
constexpr int kXcdChunk = 16;

__device__ __forceinline__ int64_t xcd_item(unsigned bid) {
    const int64_t x = bid & 7, j = bid >> 3;
    return ((j / kXcdChunk) * 8 + x) * kXcdChunk + (j % kXcdChunk);
}

int64_t xcd_grid(int64_t items);

// Grouped layout kernels over several transforms (one launch each).
struct LayoutGeom {
    const double* ao;      // packed AO block
    const double* c;       // caller's C, column-major, ld ldc
    double* vsq;           // full AO supermatrix, rows ldv apart
    double* h;             // H
    double* cr;            // zero-padded C in 16-column blocks (cr_index)
    int64_t pa, pm;
    int64_t ldv;           // row stride of vsq (pa rounded up to 16 doubles)
    int n, m, np, mp, ldc;
    int64_t expand_start;  // first expand tile of this transform
    int64_t trans_start;   // first transpose tile
    int64_t padc_start;    // first pad_c element
};
hipError_t launch_group_pad_c(const LayoutGeom* g, int ngeom, int64_t total, hipStream_t st);
hipError_t launch_group_expand(const LayoutGeom* g, int ngeom, int64_t total_tiles, hipStream_t st);
hipError_t launch_group_transpose(const LayoutGeom* g, int ngeom, int64_t total_tiles,
                                  hipStream_t st);
int64_t expand_tiles(int64_t pa);
int64_t transpose_tiles(int64_t rows, int64_t cols);

}  // namespace cipsi
