// Split-FP16 (Ozaki) half-transform on FP16 MFMA with FP32 accumulation.
//
// Each half (tpqkld.f:82-97 / tijkld.f:53-76) is run as two slice products:
//   step 1 (Q1/Q3)  T_b = Z_b C       k_split_rows, one work-group per matrix b: Z_b rows
//                                     gathered from the packed symmetric rows; stored
//                                     transposed, U[b][c][r] = T_b(r, c), with the row
//                                     exponents of U
//   step 2 (Q2/Q4)  O_b^T = U_b C     k_split_tri: rows (b, c) of U; only a <= c is kept
//                                     and written to the packed output exactly as the
//                                     FP64 kernel does.
// In both the left operand is a tall FP64 matrix A with K = N_AO <= 256 columns and the
// right operand is C.  Every row of A (and every column of C) is scaled by a power of
// two 2^e with |row| < 2^e and rounded to an integer Y of 8S-1 bits,
//   Y = rint(x 2^(8S-1-e)),  Y = sum_i d_i 256^(S-i),  d_i in [-128, 128] (balanced digits),
// so x = sum_i d_i 2^(e+1-8i) up to 2^(e-8S).  The digits are exact in FP16 and every
// digit product is at most 2^14 in magnitude, so a sum of up to four digit products
// over K <= 256 stays below 2^24 and FP32 accumulation is exact: the products with
// i + j = d are accumulated together, group d = 2..S+1 (S(S+1)/2 digit products in
// all; the groups of five and six products have a relative weight 2^-32 or less and
// their FP32 rounding is negligible).  The epilogue forms
//   out = sum_d 2^(e_row + f_col + 2 - 8d) acc_d  in FP64,
// by Horner's rule in 2^-8 (one FMA per group and one ldexp, instead of an ldexp and an
// add per group).
// The slice count S (3..6) sets the accuracy: the truncation error of an operand is
// 2^-8S of its row maximum.
//
// MFMA lane layouts of v_mfma_f32_16x16x32_f16 (l = lane, lm = l % 16, lg = l / 16):
//   A operand lane l holds A[lm][8 lg + e], B operand B[8 lg + e][lm], e = 0..7;
//   D element q of lane l is D[4 lg + q][lm].
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>

#include "../../include/cipsi_ijkl.h"
#include "common.h"
#include "ctx.h"

using namespace cipsi;

namespace {

constexpr int kMinSlices = 3;
constexpr int kMaxSlices = 6;
constexpr int kDefaultSlices = 6;
constexpr int TR = 32;        // rows of A per work item
constexpr int KCH = 32;       // K per chunk (one MFMA)
constexpr int KST = 2 * KCH;  // K per stage: the chunks staged between two barriers
constexpr int APAD = 40;      // LDS row stride of a digit plane, in halves
constexpr int MAXC = 256;     // output columns of a matrix (N_MO rounded up to 16)
// Waves per work-group: each wave owns 32 output columns.  Step 1 runs one 8-wave
// work-group per matrix over all its columns (the A digits in LDS serve 8 waves); step 2,
// whose triangle leaves many columns empty, runs two 4-wave work-groups per CU over
// column tiles of 128, so idle column tiles cost nothing and one work-group's barriers
// and epilogue overlap the other's MFMAs.
constexpr int NW1 = 8;
constexpr int NW2 = 4;
constexpr int TC = NW2 * 32;  // step-2 column tile

typedef __attribute__((address_space(1))) const double gcdouble;

__host__ __device__ inline int round_stage(int x) { return (x + KST - 1) & ~(KST - 1); }

// Step 2 keeps a <= c only: row tile rt (c < TR (rt + 1)) has tri_cols(rt) column tiles,
// and a matrix of nrt row tiles has tri_items(nrt) work items, numbered row tile by row
// tile.
__host__ __device__ inline int tri_cols(int rt) { return (TR * (rt + 1) + TC - 1) / TC; }
__host__ __device__ inline int tri_items(int nrt) {
    int t = 0;
    for (int r = 0; r < nrt; ++r) t += tri_cols(r);
    return t;
}
__device__ inline void tri_item(int p, int& rt, int& ct) {
    rt = 0;
    while (p >= tri_cols(rt)) {
        p -= tri_cols(rt);
        ++rt;
    }
    ct = p;
}

// Exponent e with |x| < 2^e (x != 0).
__device__ inline int exp_above(double x) { return ilogb(fabs(x)) + 1; }

// Balanced base-256 digits of Y = rint(x 2^(8S-1-e)), most significant first.
template <int S>
__device__ inline void digits(double x, int e, _Float16* d) {
    const double y = ldexp(x, 8 * S - 1 - e);                     // |y| <= 2^(8S-1)
    const double t = y + 6755399441055744.0;                      // 1.5 * 2^52: rint
    long long Y = (long long)__double_as_longlong(t) - 0x4338000000000000LL;
#pragma unroll
    for (int s = S - 1; s >= 1; --s) {
        int v = (int)(Y & 255);
        v -= (v & 128) << 1;                                      // [-128, 127]
        d[s] = (_Float16)(float)v;
        Y = (Y - v) >> 8;
    }
    d[0] = (_Float16)(float)(int)Y;                               // [-128, 128]
}

// The same digits for two values at once, as packed pairs of FP16: w[s] holds digit s
// of x0 in its low half and of x1 in its high half.  Adding 1.5 * 2^52 + B_S, with
// B_S = 128 (1 + 256 + ... + 256^(S-1)), rounds y to an integer and biases every
// base-256 field by 128, so that the fields of the low word bits are the balanced digits
// plus 128 (the top field, 9 bits, is d_0 + 128 in [0, 256]); a field f becomes the FP16
// value 1024 + f by setting the exponent bits 0x6400, and subtracting 1152 leaves the
// digit.  Equal to digits<S> element by element.
template <int S>
__device__ constexpr double digit_magic() {
    double b = 0.0, p = 1.0;
    for (int i = 0; i < S; ++i) {
        b += 128.0 * p;
        p *= 256.0;
    }
    return 6755399441055744.0 + b;
}

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

template <int S>
__device__ inline void digits_pair(double x0, double x1, int e0, int e1, uint32_t (&w)[S]) {
    constexpr double M = digit_magic<S>();
    const uint64_t b0 = (uint64_t)__double_as_longlong(ldexp(x0, 8 * S - 1 - e0) + M);
    const uint64_t b1 = (uint64_t)__double_as_longlong(ldexp(x1, 8 * S - 1 - e1) + M);
    const half2_t off = half2_t{(_Float16)1152.0f, (_Float16)1152.0f};
#pragma unroll
    for (int i = 0; i < S; ++i) {
        const int sh = 8 * (S - 1 - i);                     // digit i sits at bit 8 (S-1-i)
        const uint32_t mask = i == 0 ? 0x1FFu : 0xFFu;
        const uint32_t f0 = (uint32_t)(b0 >> sh) & mask, f1 = (uint32_t)(b1 >> sh) & mask;
        const uint32_t wd = f0 | (f1 << 16) | 0x64006400u;
        w[i] = __builtin_bit_cast(uint32_t, __builtin_bit_cast(half2_t, wd) - off);
    }
}

// C digits: cs[((s * kc + chunk) * mp + c) * 32 + k] for row r = 32 chunk + k, and the
// column exponents.  One thread per column.
template <int S>
__global__ void k_split_cslices(const double* __restrict__ cr, int ldm, int np, int mp, int kp,
                                _Float16* __restrict__ cs, int* __restrict__ cexp) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= mp) return;
    int e = INT_MIN;
    for (int r = 0; r < np; ++r) {
        const double x = cr[cr_index(r, c)];
        if (x != 0.0) e = max(e, exp_above(x));
    }
    if (e == INT_MIN) e = 0;
    cexp[c] = e;
    const int kc = kp / KCH;
    for (int r = 0; r < kp; ++r) {
        const double x = r < np ? cr[cr_index(r, c)] : 0.0;
        _Float16 d[S];
        digits<S>(x, e, d);
        for (int s = 0; s < S; ++s)
            cs[(((int64_t)s * kc + r / KCH) * mp + c) * KCH + (r % KCH)] = d[s];
    }
}

// Row exponents of the symmetric matrices Z_b (rows of the packed source): one
// work-group per matrix.  Each wave takes two packed rows r at a time (elements (r, s),
// s <= r, contiguous) and loads all of both rows before it reduces them, so that eight
// loads per lane are in flight; the row's maximum |x| comes from a wave reduction, and each
// element with s < r is handed to row s by an LDS atomic max on distinct addresses.
// Addresses past a row are clamped to its last element and their values dropped, so every
// load is unconditional.  Maxima are kept as the bits of |x| (ordered like the values), so
// the exponent is that of the largest element, as element-wise exponents would give.
__global__ __launch_bounds__(256) void k_split_rowexp(const double* __restrict__ src,
                                                      int64_t src_ld, int n, int kp,
                                                      int* __restrict__ rexp) {
    __shared__ unsigned long long mx[256];
    const int64_t b = blockIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    mx[threadIdx.x] = 0ull;
    __syncthreads();
    const unsigned long long* row = (const unsigned long long*)(src + b * src_ld);
    constexpr unsigned long long kAbs = 0x7FFFFFFFFFFFFFFFull;
    for (int r0 = w; r0 < n; r0 += 8) {
        unsigned long long v[2][4];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int r = r0 + 4 * q;
            const int rc = r < n ? r : n - 1;
            const unsigned long long* seg = row + pair0(rc, 0);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int s = lane + 64 * k;
                v[q][k] = seg[s <= rc ? s : rc];
            }
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int r = r0 + 4 * q;
            if (r >= n) continue;                          // the same for the whole wave
            unsigned long long m = 0ull;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int s = lane + 64 * k;
                const unsigned long long a = s <= r ? (v[q][k] & kAbs) : 0ull;
                m = a > m ? a : m;
                if (s < r) atomicMax(&mx[s], a);
            }
            for (int o = 32; o >= 1; o >>= 1) {
                const unsigned long long t = __shfl_xor(m, o);
                m = t > m ? t : m;
            }
            if (lane == 0) atomicMax(&mx[r], m);
        }
    }
    __syncthreads();
    for (int r = threadIdx.x; r < kp; r += 256) {
        const unsigned long long v = r < n ? mx[r] : 0ull;
        rexp[b * kp + r] = v == 0ull ? 0 : exp_above(__longlong_as_double((long long)v));
    }
}

struct SplitArgs {
    const double* src;       // step 1: packed symmetric rows; step 2: U rows (length kp)
    int64_t src_ld;
    const int* rexp;         // step 1: [b][kp] row exponents; step 2: [b][mp] U row exponents
    const _Float16* cs;      // C digits
    const int* cexp;         // C column exponents
    double* dst;             // step 1: U [b][mp][kp]; step 2: packed output
    int* dexp;               // step 1: U row exponents [b][mp]
    int n, m, mp, kp;
    int half;
    int64_t nb;
    int64_t dst_ld;          // step 2, half 1: P_M
    const int64_t* tile_start;  // step 2, half 2: prefix over l of (l+1) step2_items(l)
    int64_t nitems;
    int store_t;             // step 2, half 1: write H^T, element p of row b at p * nb + b
};

// Step 2: one work item is TR rows c of U_b times the TC columns a of column tile ct,
// for the (rt, ct) pairs of the triangle a <= c (tri_items per matrix).  Half 1 takes
// consecutive matrices b for one pair (H^T stores: neighbouring items, which run on one
// XCD, write neighbouring doubles of each H^T row) or consecutive pairs of one b; half 2
// finds the MO pair row l in the prefix table.  Two 4-wave work-groups run per CU.
template <int S>
__global__ __launch_bounds__(NW2 * 64, 2) void k_split_tri(SplitArgs A) {
    constexpr int NTHR = NW2 * 64;
    constexpr int EPT = TR * KCH / NTHR;           // A elements per thread and chunk
    static_assert(EPT == 4, "four A elements per thread and chunk");
    __shared__ _Float16 As[2][2][S][TR][APAD];     // [buffer][chunk of the stage][digit]
    __shared__ int rx[TR];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int lm = lane & 15, lg = lane >> 4;
    const int64_t item = xcd_item(blockIdx.x);
    if (item >= A.nitems) return;

    // ---- work item -> matrix b, row tile, column tile, extents
    int64_t b, limit, dst_off;
    int rt, ct, nrows, p;
    const int kc = A.kp / KCH;
    if (A.half == 1) {
        const int npm = tri_items((A.m + TR - 1) / TR);
        if (A.store_t) {
            p = (int)(item / A.nb);
            b = item - (int64_t)p * A.nb;
        } else {
            b = item / npm;
            p = (int)(item - b * npm);
        }
        nrows = A.m;
        limit = A.dst_ld;
        dst_off = b * A.dst_ld;
    } else {
        int lo = 0, hi = A.m;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (A.tile_start[mid] <= item) lo = mid; else hi = mid;
        }
        const int l = lo;
        const int npm = tri_items((l + TR) / TR);
        const int64_t rem = item - A.tile_start[l];
        b = pair0(l, 0) + rem / npm;
        p = (int)(rem % npm);
        nrows = l + 1;
        limit = b + 1;
        dst_off = pair0(b, 0);
    }
    tri_item(p, rt, ct);
    const int ncols = min(rt * TR + TR, nrows);                     // a <= c
    const int row0 = rt * TR;
    const int cw = ct * TC + 32 * w;                                // this wave's columns
    const bool active = cw < ncols;

    // ---- A chunk staging: this thread's EPT elements (row si, k = sk .. sk + EPT - 1)
    const int si = tid / (KCH / EPT), sk = EPT * (tid % (KCH / EPT));
    gcdouble* __restrict__ src = (gcdouble*)A.src;
    // Loads are unconditional (an element outside the matrix reads element 0 of its row,
    // which exists) and masked at commit, so no register is written by anything but its
    // load and the loop body has no branches around loads.
    auto elem = [&](int u, int ch, int64_t& idx) -> bool {
        const int r = row0 + si, s = ch * KCH + sk + u;
        const bool ok = r < nrows && s < A.n;
        idx = ok ? (b * A.mp + r) * (int64_t)A.kp + s : b * A.mp * (int64_t)A.kp;
        return ok;
    };
    auto fetch = [&](int st, double (&xs)[2 * EPT]) {
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int u = 0; u < EPT; ++u) {
                int64_t idx;
                (void)elem(u, 2 * st + c, idx);
                xs[EPT * c + u] = src[idx];
            }
    };
    auto commit = [&](int st, int buf, const double (&xs)[2 * EPT]) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            double v[EPT];
#pragma unroll
            for (int u = 0; u < EPT; ++u) {
                int64_t idx;
                v[u] = elem(u, 2 * st + c, idx) ? xs[EPT * c + u] : 0.0;
            }
            uint32_t w0[S], w1[S];
            digits_pair<S>(v[0], v[1], rx[si], rx[si], w0);
            digits_pair<S>(v[2], v[3], rx[si], rx[si], w1);
#pragma unroll
            for (int s = 0; s < S; ++s) *(uint2*)&As[buf][c][s][si][sk] = make_uint2(w0[s], w1[s]);
        }
    };

    f32x4 acc[S][2][2];
#pragma unroll
    for (int g = 0; g < S; ++g)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[g][i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // B fragments (digits of C), double-buffered in registers: the next chunk's S x 2
    // fragments are loaded at the start of a chunk, ahead of its A prefetch, by every
    // wave (a wave without columns reads a clamped column), so that no load sits in a
    // branch.  Loads complete in order: the commit waits only for the A prefetch, and a
    // chunk's MFMAs wait only for B loads issued a whole chunk earlier.  Columns past mp
    // only feed outputs that are never stored.  The A fragments are read from LDS per
    // digit i, so only one pair is live.
    // The loads go through a buffer resource: the lane's column offset is one VGPR per
    // column block, and the digit plane and chunk a scalar offset.
    f16x8 bcur[S][2], bnext[S][2];
    const __amdgpu_buffer_rsrc_t crs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)A.cs, (short)0, (int)(sizeof(_Float16) * (size_t)S * A.kp * A.mp), 0x00020000);
    unsigned boff[2];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
        boff[cb] = (unsigned)(min(cw + 16 * cb + lm, A.mp - 1) * KCH + 8 * lg) * 2u;
    auto load_b = [&](int ch, f16x8 (&bf)[S][2]) {
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const unsigned so = (unsigned)(((j * kc + ch) * A.mp) * KCH) * 2u;
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
                bf[j][cb] = __builtin_bit_cast(
                    f16x8, __builtin_amdgcn_raw_buffer_load_b128(crs, boff[cb], so, 0));
        }
    };
    auto mfma_chunk = [&](int buf, int c, const f16x8 (&bc)[S][2]) {
#pragma unroll
        for (int i = 0; i < S; ++i) {
            f16x8 af[2];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
                af[rb] = *(const f16x8*)&As[buf][c][i][16 * rb + lm][8 * lg];
#pragma unroll
            for (int j = 0; i + j < S; ++j)
#pragma unroll
                for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                    for (int cb = 0; cb < 2; ++cb)
                        acc[i + j][rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                            af[rb], bc[j][cb], acc[i + j][rb][cb], 0, 0, 0);
        }
    };
    // K runs in stages of two chunks, one barrier per stage.  The A elements are fetched
    // two stages ahead (their rows mostly come from HBM), the B fragments one chunk ahead
    // (they stay in L2).  Fetches past the last stage or chunk are clamped to it, so that
    // no load is conditional.
    const int nst = kc / 2;                        // kp is a whole number of stages
    double xa[2 * EPT], xb[2 * EPT];
    // The first loads go out before the row exponents are read, so that their latencies
    // overlap.
    load_b(0, bcur);
    fetch(0, xa);
    fetch(nst > 1 ? 1 : 0, xb);
    if (tid < TR) {
        const int r = row0 + tid;
        int e = 0;
        if (r < nrows) {
            e = A.rexp[b * A.mp + r];
            if (e == INT_MIN) e = 0;
        }
        rx[tid] = e;
    }
    __syncthreads();
    commit(0, 0, xa);
    __syncthreads();
    // The two B buffers alternate between the two chunks of a stage: bcur feeds the even
    // chunk and is reloaded for the next stage as soon as its MFMAs are issued.  Loads
    // complete in order, so the A prefetch is issued after that reload: the odd chunk's
    // MFMAs wait for B loads issued before it and not for the HBM gathers, which have
    // until the next stage's odd chunk.
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1;
        load_b(2 * st + 1, bnext);
        __builtin_amdgcn_sched_barrier(0);
        if (active) mfma_chunk(buf, 0, bcur);
        __builtin_amdgcn_sched_barrier(0);
        load_b(min(2 * st + 2, kc - 1), bcur);
        fetch(min(st + 2, nst - 1), xa);
        __builtin_amdgcn_sched_barrier(0);
        if (active) mfma_chunk(buf, 1, bnext);
        __builtin_amdgcn_sched_barrier(0);
        if (st + 1 < nst) commit(st + 1, buf ^ 1, xb);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 2 * EPT; ++u) xb[u] = xa[u];
    }
    if (!active) return;

    // ---- epilogue: FP64 recombination and stores into the packed output
    typedef __attribute__((address_space(1))) double gdouble;
    gdouble* __restrict__ dst = (gdouble*)A.dst;
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        const int col = cw + 16 * cb + lm;
        const int fc = col < A.mp ? A.cexp[col] : 0;
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int il = 16 * rb + 4 * lg + q;
                const int base = rx[il] + fc + 2;
                double v = (double)acc[S - 1][rb][cb][q];
#pragma unroll
                for (int g = S - 2; g >= 0; --g)
                    v = fma(v, 0.00390625, (double)acc[g][rb][cb][q]);
                v = ldexp(v, base - 16);
                const int r = row0 + il;
                if (r < nrows && col <= r) {
                    const int64_t p = pair0(r, col);
                    if (p < limit) {
                        if (A.store_t) dst[p * A.nb + b] = v;
                        else dst[dst_off + p] = v;
                    }
                }
            }
        }
    }
}

// Step 1 as one work-group per matrix b: the row tiles of Z_b run as one pipeline of
// K stages (tile t, stage k is global stage t * nst + k), so the A gathers of the next
// tile are in flight while a tile's epilogue runs, and the work-group starts only once
// per matrix.  Each wave owns 32 columns over all rows of b, so the exponent of U row
// (b, col), the maximum over every row tile, is reduced in registers and written once.
// Row exponents of the A rows are loaded one tile ahead into registers: the committing
// thread's row (rc0 / rc1 by tile parity) and the 8 rows of the lane's output elements.
template <int S>
__global__ __launch_bounds__(NW1 * 64, 1) void k_split_rows(SplitArgs A) {
    constexpr int NTHR = NW1 * 64;
    constexpr int EPT = TR * KCH / NTHR;
    static_assert(EPT == 2, "two A elements per thread and chunk");
    __shared__ _Float16 As[2][2][S][TR][APAD];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int lm = lane & 15, lg = lane >> 4;
    const int64_t b = xcd_item(blockIdx.x);
    if (b >= A.nitems) return;
    const int kc = A.kp / KCH, nst = kc / 2;
    const int ntr = (A.n + TR - 1) / TR;
    const int ntot = ntr * nst;                                    // stages of the matrix
    const int ncols = A.half == 1 ? A.m : (int)pair_row(b) + 1;   // half 2: columns j <= l
    const int cw = 32 * w;
    const bool active = cw < ncols;

    const int si = tid / (KCH / EPT), sk = EPT * (tid % (KCH / EPT));
    gcdouble* __restrict__ src = (gcdouble*)A.src + b * A.src_ld;
    const int* __restrict__ rexp = A.rexp + b * A.kp;
    // A element u of chunk ch of global stage g: row t * TR + si of Z_b, column s.
    auto elem = [&](int u, int g, int ch, int64_t& idx) -> bool {
        const int t = g / nst;
        const int r = t * TR + si, s = (g - t * nst) * KST + ch * KCH + sk + u;
        const bool ok = r < A.n && s < A.n;
        const int hi = r >= s ? r : s, lo = r >= s ? s : r;
        idx = ok ? pair0(hi, lo) : 0;
        return ok;
    };
    auto fetch = [&](int g, double (&xs)[2 * EPT]) {
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int u = 0; u < EPT; ++u) {
                int64_t idx;
                (void)elem(u, g, c, idx);
                xs[EPT * c + u] = src[idx];
            }
    };
    auto commit = [&](int g, int buf, int e, const double (&xs)[2 * EPT]) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            int64_t idx;
            const double v0 = elem(0, g, c, idx) ? xs[EPT * c] : 0.0;
            const double v1 = elem(1, g, c, idx) ? xs[EPT * c + 1] : 0.0;
            uint32_t wd[S];
            digits_pair<S>(v0, v1, e, e, wd);
#pragma unroll
            for (int s = 0; s < S; ++s) *(uint32_t*)&As[buf][c][s][si][sk] = wd[s];
        }
    };
    auto row_exp = [&](int r) { return r < A.n ? rexp[r] : 0; };

    f32x4 acc[S][2][2];
    auto clear = [&]() {
#pragma unroll
        for (int g = 0; g < S; ++g)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[g][i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    };
    clear();

    f16x8 bcur[S][2], bnext[S][2];
    const __amdgpu_buffer_rsrc_t crs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)A.cs, (short)0, (int)(sizeof(_Float16) * (size_t)S * A.kp * A.mp), 0x00020000);
    unsigned boff[2];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
        boff[cb] = (unsigned)(min(cw + 16 * cb + lm, A.mp - 1) * KCH + 8 * lg) * 2u;
    auto load_b = [&](int ch, f16x8 (&bf)[S][2]) {
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const unsigned so = (unsigned)(((j * kc + ch) * A.mp) * KCH) * 2u;
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
                bf[j][cb] = __builtin_bit_cast(
                    f16x8, __builtin_amdgcn_raw_buffer_load_b128(crs, boff[cb], so, 0));
        }
    };
    auto mfma_chunk = [&](int buf, int c, const f16x8 (&bc)[S][2]) {
#pragma unroll
        for (int i = 0; i < S; ++i) {
            f16x8 af[2];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
                af[rb] = *(const f16x8*)&As[buf][c][i][16 * rb + lm][8 * lg];
#pragma unroll
            for (int j = 0; i + j < S; ++j)
#pragma unroll
                for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                    for (int cb = 0; cb < 2; ++cb)
                        acc[i + j][rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                            af[rb], bc[j][cb], acc[i + j][rb][cb], 0, 0, 0);
        }
    };

    int fc[2];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        const int col = cw + 16 * cb + lm;
        fc[cb] = col < A.mp ? A.cexp[col] : 0;
    }
    int emax[2] = {INT_MIN, INT_MIN};
    typedef __attribute__((address_space(1))) double gdouble;
    gdouble* __restrict__ dst = (gdouble*)A.dst;

    double xa[2 * EPT], xb[2 * EPT];
    load_b(0, bcur);
    fetch(0, xa);
    fetch(ntot > 1 ? 1 : 0, xb);
    int rc0 = row_exp(si), rc1 = row_exp(TR + si);                 // committing rows, tiles 0, 1
    int er[2][4];                                                  // output rows of this tile
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int q = 0; q < 4; ++q) er[rb][q] = row_exp(16 * rb + 4 * lg + q);
    commit(0, 0, rc0, xa);
    __syncthreads();
    for (int g = 0; g < ntot; ++g) {
        const int buf = g & 1;
        const int t = g / nst, k = g - t * nst;
        load_b(2 * k + 1, bnext);
        __builtin_amdgcn_sched_barrier(0);
        if (active) mfma_chunk(buf, 0, bcur);
        __builtin_amdgcn_sched_barrier(0);
        load_b(2 * k + 2 < kc ? 2 * k + 2 : 0, bcur);
        fetch(min(g + 2, ntot - 1), xa);
        __builtin_amdgcn_sched_barrier(0);
        if (active) mfma_chunk(buf, 1, bnext);
        __builtin_amdgcn_sched_barrier(0);
        if (g + 1 < ntot) {
            const int t1 = (g + 1) / nst;
            commit(g + 1, buf ^ 1, (t1 & 1) ? rc1 : rc0, xb);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 2 * EPT; ++u) xb[u] = xa[u];
        if (k == nst - 1) {
            // ---- tile t done: FP64 recombination into U rows (b, col), element r
            const int row0 = t * TR;
            if (active) {
#pragma unroll
                for (int cb = 0; cb < 2; ++cb) {
                    const int col = cw + 16 * cb + lm;
#pragma unroll
                    for (int rb = 0; rb < 2; ++rb) {
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int il = 16 * rb + 4 * lg + q;
                            const int base = er[rb][q] + fc[cb] + 2;
                            double v = (double)acc[S - 1][rb][cb][q];
#pragma unroll
                            for (int gg = S - 2; gg >= 0; --gg)
                                v = fma(v, 0.00390625, (double)acc[gg][rb][cb][q]);
                            v = ldexp(v, base - 16);
                            const int r = row0 + il;
                            if (r < A.n && col < ncols) {
                                dst[(b * A.mp + col) * (int64_t)A.kp + r] = v;
                                if (v != 0.0) emax[cb] = max(emax[cb], exp_above(v));
                            }
                        }
                    }
                }
            }
            clear();
            // exponents of the tile after next (committing rows) and of the next tile
            // (output rows)
            if (t & 1) rc1 = row_exp((t + 2) * TR + si);
            else rc0 = row_exp((t + 2) * TR + si);
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int q = 0; q < 4; ++q) er[rb][q] = row_exp((t + 1) * TR + 16 * rb + 4 * lg + q);
        }
    }
    // U row exponents: maximum over the four row groups of the wave's lanes
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        int e = emax[cb];
        e = max(e, __shfl_xor(e, 16));
        e = max(e, __shfl_xor(e, 32));
        const int col = cw + 16 * cb + lm;
        if (lg == 0 && col < ncols) A.dexp[b * A.mp + col] = e;
    }
}

template <int S>
hipError_t launch_split_rows(const SplitArgs& a, hipStream_t st) {
    const int64_t g = xcd_grid(a.nitems);
    if (g <= 0 || g > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_split_rows<S>), dim3((unsigned)g), dim3(NW1 * 64), 0, st, a);
    return hipGetLastError();
}

template <int S>
hipError_t launch_split_tri(const SplitArgs& a, hipStream_t st) {
    const int64_t g = xcd_grid(a.nitems);
    if (g <= 0 || g > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_split_tri<S>), dim3((unsigned)g), dim3(NW2 * 64), 0, st, a);
    return hipGetLastError();
}

template <int STEP>
hipError_t launch_split_s(int s, const SplitArgs& a, hipStream_t st) {
    switch (s) {
        case 3: return STEP == 1 ? launch_split_rows<3>(a, st) : launch_split_tri<3>(a, st);
        case 4: return STEP == 1 ? launch_split_rows<4>(a, st) : launch_split_tri<4>(a, st);
        case 5: return STEP == 1 ? launch_split_rows<5>(a, st) : launch_split_tri<5>(a, st);
        case 6: return STEP == 1 ? launch_split_rows<6>(a, st) : launch_split_tri<6>(a, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_cslices(int s, const double* cr, int ldm, int np, int mp, int kp, _Float16* cs,
                          int* cexp, hipStream_t st) {
    const dim3 grid((mp + 63) / 64), block(64);
    switch (s) {
        case 3: hipLaunchKernelGGL(k_split_cslices<3>, grid, block, 0, st, cr, ldm, np, mp, kp, cs, cexp); break;
        case 4: hipLaunchKernelGGL(k_split_cslices<4>, grid, block, 0, st, cr, ldm, np, mp, kp, cs, cexp); break;
        case 5: hipLaunchKernelGGL(k_split_cslices<5>, grid, block, 0, st, cr, ldm, np, mp, kp, cs, cexp); break;
        case 6: hipLaunchKernelGGL(k_split_cslices<6>, grid, block, 0, st, cr, ldm, np, mp, kp, cs, cexp); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct SplitLayout {
    size_t u, uexp, rexp, cs, cexp, total;
};

SplitLayout split_layout(int n, int m, int s) {
    const int64_t pa = npairs(n), pm = npairs(m);
    const int64_t nb = pa > pm ? pa : pm;
    const int kp = round_stage(n), mp = (m + 15) & ~15;
    SplitLayout L;
    size_t off = 0;
    L.u = off;
    off = align256(off + sizeof(double) * nb * mp * kp);
    L.uexp = off;
    off = align256(off + sizeof(int) * nb * mp);
    L.rexp = off;
    off = align256(off + sizeof(int) * nb * kp);
    L.cs = off;
    off = align256(off + sizeof(_Float16) * (size_t)s * kp * mp);
    L.cexp = off;
    off = align256(off + sizeof(int) * mp);
    L.total = off;
    return L;
}

}  // namespace

bool split_supported(int n_slices) {
    return n_slices == 0 || (n_slices >= kMinSlices && n_slices <= kMaxSlices);
}

int64_t split_workspace_bytes(int n_ao, int n_mo) {
    return (int64_t)split_layout(n_ao, n_mo, kMaxSlices).total;
}

// Unit of the a-priori error estimate: the truncation of the two operands of a slice
// product, 2^-8S of their row (column) maxima each.
double split_unit_roundoff(int, int n_slices) {
    const int s = n_slices == 0 ? kDefaultSlices : n_slices;
    return std::ldexp(1.0, 1 - 8 * s);
}

// Step-2 work items of one MO pair of row l in half 2 (its rows c <= l).
static int64_t step2_items(int l, int) { return tri_items((l + TR) / TR); }

int run_half_split(cipsi_ijkl_ctx* ctx, const HalfArgs& a, int, int n_slices) {
    const int s = n_slices == 0 ? kDefaultSlices : n_slices;
    if (s < kMinSlices || s > kMaxSlices) return CIPSI_ERR_UNSUPPORTED;
    const int n = a.n, m = a.m, mp = a.mp, kp = round_stage(n);
    if (mp > MAXC || kp > 256) return CIPSI_ERR_UNSUPPORTED;
    const SplitLayout L = split_layout(n, m, kMaxSlices);
    int rc = ctx_alloc(ctx, ctx->slices, L.total);
    if (rc) return rc;
    char* base = (char*)ctx->slices.ptr;
    double* u = (double*)(base + L.u);
    int* uexp = (int*)(base + L.uexp);
    int* rexp = (int*)(base + L.rexp);
    _Float16* cs = (_Float16*)(base + L.cs);
    int* cexp = (int*)(base + L.cexp);
    hipStream_t st = ctx->stream;
    hipError_t e;
    if ((e = launch_cslices(s, a.cr, a.ldm, a.np, mp, kp, cs, cexp, st)) != hipSuccess)
        return CIPSI_ERR_HIP;
    hipLaunchKernelGGL(k_split_rowexp, dim3((unsigned)a.nb), dim3(256), 0, st, a.src, a.src_ld, n,
                       kp, rexp);
    if (hipGetLastError() != hipSuccess) return CIPSI_ERR_HIP;

    SplitArgs x{};
    x.cs = cs;
    x.cexp = cexp;
    x.n = n;
    x.m = m;
    x.mp = mp;
    x.kp = kp;
    x.half = a.half;
    x.nb = a.nb;
    // step 1: T_b = Z_b C, stored transposed into U
    x.src = a.src;
    x.src_ld = a.src_ld;
    x.rexp = rexp;
    x.dst = u;
    x.dexp = uexp;
    x.nitems = a.nb;
    if ((e = launch_split_s<1>(s, x, st)) != hipSuccess) return CIPSI_ERR_HIP;
    // step 2: O_b^T = U_b C, packed triangle
    x.src = u;
    x.src_ld = 0;
    x.rexp = uexp;
    x.dst = a.dst;
    x.dexp = nullptr;
    if (a.half == 1) {
        x.dst_ld = a.dst_ld;
        x.store_t = a.store_t;
        x.tile_start = nullptr;
        x.nitems = a.nb * (int64_t)tri_items((m + TR - 1) / TR);
    } else {
        const int64_t* ts;
        int64_t total;
        if ((rc = ctx_prefix_table(ctx, std::make_pair(m, -1), m, step2_items, &ts, &total))) return rc;
        x.tile_start = ts;
        x.nitems = total;
    }
    if ((e = launch_split_s<2>(s, x, st)) != hipSuccess) return CIPSI_ERR_HIP;
    return CIPSI_OK;
}
