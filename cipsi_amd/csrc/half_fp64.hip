// FP64 MFMA half-transform ("sandwich") kernel for gfx950.
//
// One work-group = one symmetric matrix Z_b (an AO pair for half 1, an MO pair for
// half 2) times one tile of up to NW 16-column blocks; the nt tiles of a matrix take
// the column blocks round robin (tile t: blocks t, t + nt, t + 2nt, ...), so every tile
// has work for about the same number of waves and a similar share of the triangle.
// Wave w owns one block, 16 columns c_w .. c_w + 15, and every row block:
//   Q1/Q3:  T(r, c) = sum_s Z_b(r, s) C(s, c)          (tpqkld.f:82-90 / tijkld.f:53-60)
//   Q2/Q4:  O(a, c) = sum_r C(r, a) T(r, c), a <= c    (tpqkld.f:91-97 / tijkld.f:61-76)
//
// v_mfma_f64_16x16x4f64 lane layouts (l = lane, lm = l%16, lg = l/16):
//   A[m][k] = lane (m, k) -> A[lm][lg];  B[k][n] -> B[lg][lm];
//   D element q of lane l  = D[lg + 4q][lm].
// The D layout of a T row block is therefore exactly the operand layout of the four
// K-steps r = 16rb + 4q + {0..3}: T never leaves the registers.  Q2 is computed as
// O^T(c, a) = sum_r T(r, c) C(r, a) with T as the A operand, so a lane's results are
// 16 consecutive a of one c -- 128-byte coalesced stores into the packed output.  Half 1
// may instead store H^T (store_t): element (a, c) of matrix b at pair0(c, a) * nb + b, one
// 128-byte line per lane, completed in L2 by the neighbouring AO pairs of the same line.
//
// Only the Z_b panels (KC columns s of the symmetric matrix, all rows) go through LDS,
// prefetched into registers one panel ahead; the C operands of both products are read
// from the zero-padded copies of C in global memory (L2-resident): the first product
// from cr_index order, the second from cq_index order.
#include "common.h"

#include <utility>
#include <vector>

namespace cipsi {

template <int NW, int NRB, int KC>
struct HalfCfg {
    static constexpr int NT = NW * 64;
    static constexpr int LT = NW * 16;                       // columns per work-group
    static constexpr int EP = 16 * KC;                       // panel elements per 16-row pass
    static constexpr int PPS = NT >= EP ? NT / EP : 1;       // passes per sweep of the threads
    static constexpr int SUB = NT >= EP ? 1 : EP / NT;       // sweeps per pass
    static constexpr int NU = ((NRB + PPS - 1) / PPS) * SUB; // loads per thread per panel
};

// Global-memory views of the argument pointers: when the arguments come from a device
// table (grouped launch) the compiler cannot tell that they point to global memory and
// would emit flat accesses.
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) const f64x2 gcdouble2;
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) const int64_t gcint64;

// The value the Q1 staging loads for panel elements outside the matrix.
__device__ double kZeroElement[1] = {0.0};

__device__ inline void work_item(const HalfArgs& A, int lt, int64_t item, int64_t& b, int& tile,
                                 int& ncols, int64_t& limit, int64_t& dst_off) {
    if (A.half == 1) {
        const int nt = (A.m + lt - 1) / lt;
        b = A.b_begin + item / nt;
        tile = (int)(item % nt);
        ncols = A.m;
        limit = A.dst_ld;
        dst_off = b * A.dst_ld;
    } else {
        int lo = 0, hi = A.m;                      // tile_start[0] = 0 <= item < tile_start[m]
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (((gcint64*)A.tile_start)[mid] <= item) lo = mid; else hi = mid;
        }
        const int l = lo;
        const int nt = (l + lt) / lt;              // ceil((l+1)/lt)
        const int64_t rem = item - ((gcint64*)A.tile_start)[l];
        b = pair0(l, 0) + rem / nt;
        tile = (int)(rem % nt);
        ncols = l + 1;
        limit = b + 1;
        dst_off = pair0(b, 0);
    }
}

// One work item (matrix b, column tile) of one half.  Every thread of the work-group
// calls it with the same item; it ends with a barrier, so LDS is free on return.
//
// Only the Z_b panels go through LDS (they are shared by all waves).  The C operands
// are read from global memory (C is a few hundred KB and stays in L2): in Q1 one value
// per lane per 4-deep K step serves all NRB row blocks; in Q2 every wave reads its own
// C blocks, so Q2 needs no LDS and no barriers and each wave works through its own
// triangle (column block k does k + 1 steps) at its own pace.
template <int NW, int NRB, int KC>
__device__ __forceinline__ void half_tile(const HalfArgs& A, int64_t item) {
    using K = HalfCfg<NW, NRB, KC>;
    extern __shared__ double smem[];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int lm = lane & 15, lg = lane >> 4;
    const int np = A.np, n = A.n;
    constexpr int RP = NRB * 16;                   // rows held in LDS (zero beyond n)
    constexpr int LDZ = RP + 1;                    // odd: conflict-free panel commits

    int64_t b, limit, dst_off;
    int tile, ncols;
    work_item(A, K::LT, item, b, tile, ncols, limit, dst_off);
    const int ntiles = (ncols + K::LT - 1) / K::LT;
    const int c1 = ncols;
    // Column block of this wave: slot wc of the tile, block tile + ntiles * wc.  The
    // tile has nact busy slots, and slot i's Q2 work grows with i.  Waves w and w + 4
    // share SIMD w % 4, so with 8 waves P = max(0, nact - 4) SIMDs must hold two busy
    // waves: SIMD j < P takes slots j and 2P - 1 - j (smallest with largest), the other
    // SIMDs take the largest slots 2P .. nact - 1 alone.  This evens out both the Q1
    // waves and the Q2 triangle per SIMD.  Slots >= nact are idle.
    int wc = w;
    if (NW == 8) {
        const int nblk = (ncols + 15) >> 4;
        const int nact = (nblk - tile + ntiles - 1) / ntiles;
        const int p2 = nact > 4 ? nact - 4 : 0;
        const int j = w & 3;
        if (j < p2) wc = (w >> 2) ? 2 * p2 - 1 - j : j;
        else wc = (w >> 2) ? nact + j : p2 + j;
    }
    const int cw = 16 * (tile + ntiles * wc);     // first column of this wave
    const bool wave_active = cw < c1;
    gcdouble* __restrict__ zsrc = (gcdouble*)A.src + b * A.src_ld;
    gcdouble* __restrict__ cr = (gcdouble*)A.cr;

    constexpr int PANEL = KC * LDZ;                // doubles per Q1 stage buffer

    // ---------------------------------------------------------------- Q1 staging
    // Panel = Z(r, s0 + kk) for r < np, kk < KC, taken in 16-row passes.  Below the
    // diagonal block (r >= s) a row's KC entries are contiguous in the packed row
    // (kk fastest across lanes); above it (r < s0) the entries of one s are contiguous
    // in r (r fastest across lanes).  Branch-free: an element outside the matrix is
    // loaded from a zero in global memory, so the commit stores what arrived.
    double zst[K::NU];
    auto zslot = [&](int u, int& p, int& e) {
        const int sweep = u / K::SUB, sub = u % K::SUB;
        const int t = tid + sub * K::NT;
        p = sweep * K::PPS + t / K::EP;
        e = t % K::EP;
    };
    auto zmap = [&](int p, int e, int s0, int& r, int& kk) {
        const bool above = 16 * p + 15 < s0;
        r = 16 * p + (above ? (e & 15) : e / KC);
        kk = above ? (e >> 4) : (e % KC);
    };
    auto fetch = [&](int s0) {
#pragma unroll
        for (int u = 0; u < K::NU; ++u) {
            int p, e, r, kk;
            zslot(u, p, e);
            zmap(p, e, s0, r, kk);
            const int s = s0 + kk;
            const unsigned hi = r >= s ? r : s, lo = r >= s ? s : r;
            const unsigned idx = ((hi * (hi + 1)) >> 1) + lo;
            gcdouble* ptr = (r < n && s < n) ? zsrc + idx : (gcdouble*)kZeroElement;
            zst[u] = *ptr;                 // waited for at commit, not here
        }
    };
    auto commit = [&](int s0, int boff) {
#pragma unroll
        for (int u = 0; u < K::NU; ++u) {
            int p, e, r, kk;
            zslot(u, p, e);
            zmap(p, e, s0, r, kk);
            if (p < NRB) smem[boff + kk * LDZ + r] = zst[u];
        }
    };

    // ---------------------------------------------------------------- Q1: T in registers
    f64x4 acc[NRB];
#pragma unroll
    for (int i = 0; i < NRB; ++i) acc[i] = f64x4{0.0, 0.0, 0.0, 0.0};

    // B operand C(s0 + k4 + lg, cw + lm) (C is zero-padded to kCRows rows, so every
    // chunk's rows exist).  A chunk's KC/4 values are loaded during the previous chunk,
    // before that chunk's Z prefetch: loads complete in order, so the MFMAs of a chunk
    // wait only for C loads issued a whole chunk earlier, never for the HBM prefetch of
    // the next Z panel.
    constexpr int NB = KC / 4;
    gcdouble* __restrict__ cbw = cr + (int64_t)(cw >> 4) * (kCRows * 16) + lg * 16 + lm;
    double bcur[NB];
    const int nchunk = (np + KC - 1) / KC;
    if (wave_active) {
#pragma unroll
        for (int k = 0; k < NB; ++k) bcur[k] = cbw[(4 * k) * 16];
    }
    fetch(0);
    commit(0, 0);
    __builtin_amdgcn_s_waitcnt(0x0F70);            // vmcnt(0), as at the end of a chunk
    __syncthreads();
    for (int ch = 0; ch < nchunk; ++ch) {
        const bool more = (ch + 1) < nchunk;
        double bnext[NB];
        if (more && wave_active) {
#pragma unroll
            for (int k = 0; k < NB; ++k) bnext[k] = cbw[((ch + 1) * KC + 4 * k) * 16];
        }
        __builtin_amdgcn_sched_barrier(0);
        if (more) fetch((ch + 1) * KC);
        __builtin_amdgcn_sched_barrier(0);
        if (wave_active) {
            const int zo = (ch & 1) * PANEL;
#pragma unroll
            for (int k = 0; k < NB; ++k) {
#pragma unroll
                for (int rb = 0; rb < NRB; ++rb) {
                    const double af = smem[zo + (4 * k + lg) * LDZ + rb * 16 + lm];
                    acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, bcur[k], acc[rb], 0, 0, 0);
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        if (more) commit((ch + 1) * KC, ((ch + 1) & 1) * PANEL);
        // vmcnt(0): every load of this chunk has been consumed on the paths that use it;
        // without the explicit wait the compiler keeps the skipped commit slots pending
        // into the next chunk and stalls its prefetch on them.
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        if (more && wave_active) {
#pragma unroll
            for (int k = 0; k < NB; ++k) bcur[k] = bnext[k];
        }
    }
    if (!wave_active) return;

    // ---------------------------------------------------------------- Q2: O^T = T^T C
    // K step (rb, q) covers rows r = 16 rb + 4 q + {0..3}; lane (lm, lg) holds
    // T(16 rb + 4 q + lg, cw + lm) in acc[rb][q] (the A operand) and loads
    // C(16 rb + 4 q + lg, 16 ab + lm) as the B operand.  Wave w needs a <= its last
    // column, so it stops after block cw / 16.
    const int my_last_ab = cw / 16;
    gdouble* __restrict__ dst = (gdouble*)A.dst + dst_off;
    // C rows of row block rb of column block ab (zero-padded to kCRows rows), prefetched
    // QD row blocks ahead through a ring of QD + 1.  They come from the second copy of the
    // padded C (cq_index), where a lane's four values of a row block are two 16-byte
    // pieces: two loads per row block instead of four, which measured 1.4 ms less in half 1
    // at N = 240.  The first QD row blocks of the next column block are loaded before
    // this block's results are stored: stores count in the same in-order counter as
    // loads, so loads issued after a store could only be waited for together with it.
    constexpr int QD = KC == 16 ? 3 : 1;               // registers allow a deeper ring at KC 16
    double ring[QD + 1][4];
    gcdouble* __restrict__ cq = cr + (int64_t)kCRows * A.ldm;
    auto load_rb = [&](int ab, int rb, double* bq) {
        gcdouble2* __restrict__ p = (gcdouble2*)(cq + ((int64_t)ab * 16 + rb) * 256) + lane;
        const f64x2 v0 = p[0], v1 = p[64];
        bq[0] = v0.x;
        bq[1] = v0.y;
        bq[2] = v1.x;
        bq[3] = v1.y;
    };
#pragma unroll
    for (int rb = 0; rb < QD && rb < NRB; ++rb) load_rb(0, rb, ring[rb]);
    for (int ab = 0; ab <= my_last_ab; ++ab) {
        f64x4 o0 = f64x4{0.0, 0.0, 0.0, 0.0}, o1 = o0;
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb) {
            if (rb + QD < NRB) load_rb(ab, rb + QD, ring[(rb + QD) % (QD + 1)]);
            const double* bq = ring[rb % (QD + 1)];
            o0 = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[rb][0], bq[0], o0, 0, 0, 0);
            o1 = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[rb][1], bq[1], o1, 0, 0, 0);
            o0 = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[rb][2], bq[2], o0, 0, 0, 0);
            o1 = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[rb][3], bq[3], o1, 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (ab < my_last_ab) {
#pragma unroll
            for (int rb = 0; rb < QD && rb < NRB; ++rb) load_rb(ab + 1, rb, ring[rb]);
        }
        __builtin_amdgcn_sched_barrier(0);
        // lane (lm, lg), element q: O(a = 16 ab + lm, c = cw + lg + 4q)
        const int a = ab * 16 + lm;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = cw + lg + 4 * q;
            if (c < c1 && a <= c) {
                const int64_t pp = pair0(c, a);
                if (pp < limit) {
                    if (A.store_t) ((gdouble*)A.dst)[pp * A.nb + b] = o0[q] + o1[q];
                    else dst[pp] = o0[q] + o1[q];
                }
            }
        }
    }
}

int64_t xcd_grid(int64_t items) {
    constexpr int64_t q = 8 * kXcdChunk;
    return (items + q - 1) / q * q;
}

// ---------------------------------------------------------------- half 2, paired pieces
// Half 2 of one transform with the column blocks of its matrices regrouped (8 waves, panel
// depth 16).  An MO pair of row l has nb = ceil((l + 1) / 16) column blocks.  A matrix with
// more than 8 blocks gives one full piece of 8 blocks and a remainder of nb - 8; a matrix with
// fewer than 8 gives one partial piece.  Two pieces whose blocks add up to at most 8 share a
// work-group: each is staged into its own pair of LDS panel buffers, and every wave works on
// one block of one piece, with Q1 and Q2 exactly as in half_tile (the same operations in the
// same order, so the results are the same bits).  A work-group's time depends little on how
// many of its waves are busy, so fewer, fuller work-groups finish half 2 sooner.
// The remainder of a large matrix takes the middle blocks [4, 4 + r): the full piece keeps
// blocks 0..3 and nb - 4..nb - 1, whose Q2 triangles pair up evenly on the SIMDs.
template <int NRB>
__device__ __forceinline__ void pair_tile(const HalfArgs& A, const PairItem& it) {
    constexpr int NW = 8, KC = 16;
    using K = HalfCfg<NW, NRB, KC>;
    extern __shared__ double smem[];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int lm = lane & 15, lg = lane >> 4;
    const int np = A.np, n = A.n;
    constexpr int RP = NRB * 16;
    constexpr int LDZ = RP + 1;
    constexpr int PANEL = KC * LDZ;
    const bool two = it.b[1] >= 0;
    const unsigned slot = (unsigned)(it.slots >> (8 * w)) & 0xFFu;
    const bool wave_active = (slot & 0x80u) != 0;
    const int piece = (int)((slot >> 4) & 1u);
    const int cw = 16 * (int)(slot & 15u);             // first column of this wave
    const int64_t b = it.b[piece];                     // the wave's matrix (an MO pair)
    const int64_t l = wave_active ? pair_row(b) : 0;
    const int c1 = (int)l + 1;                         // columns j <= l
    const int64_t limit = b + 1, dst_off = pair0(b, 0);
    gcdouble* __restrict__ cr = (gcdouble*)A.cr;
    gcdouble* __restrict__ zsrc0 = (gcdouble*)A.src + it.b[0] * A.src_ld;
    gcdouble* __restrict__ zsrc1 = (gcdouble*)A.src + (two ? it.b[1] : it.b[0]) * A.src_ld;

    double zst0[K::NU], zst1[K::NU];
    auto zslot = [&](int u, int& p, int& e) {
        const int sweep = u / K::SUB, sub = u % K::SUB;
        const int t = tid + sub * K::NT;
        p = sweep * K::PPS + t / K::EP;
        e = t % K::EP;
    };
    auto zmap = [&](int p, int e, int s0, int& r, int& kk) {
        const bool above = 16 * p + 15 < s0;
        r = 16 * p + (above ? (e & 15) : e / KC);
        kk = above ? (e >> 4) : (e % KC);
    };
    auto fetch = [&](int s0, gcdouble* __restrict__ zsrc, double (&zst)[K::NU]) {
#pragma unroll
        for (int u = 0; u < K::NU; ++u) {
            int p, e, r, kk;
            zslot(u, p, e);
            zmap(p, e, s0, r, kk);
            const int s = s0 + kk;
            const unsigned hi = r >= s ? r : s, lo = r >= s ? s : r;
            const unsigned idx = ((hi * (hi + 1)) >> 1) + lo;
            gcdouble* ptr = (r < n && s < n) ? zsrc + idx : (gcdouble*)kZeroElement;
            zst[u] = *ptr;
        }
    };
    auto commit = [&](int s0, int boff, const double (&zst)[K::NU]) {
#pragma unroll
        for (int u = 0; u < K::NU; ++u) {
            int p, e, r, kk;
            zslot(u, p, e);
            zmap(p, e, s0, r, kk);
            if (p < NRB) smem[boff + kk * LDZ + r] = zst[u];
        }
    };

    f64x4 acc[NRB];
#pragma unroll
    for (int i = 0; i < NRB; ++i) acc[i] = f64x4{0.0, 0.0, 0.0, 0.0};
    constexpr int NB = KC / 4;
    gcdouble* __restrict__ cbw = cr + (int64_t)(cw >> 4) * (kCRows * 16) + lg * 16 + lm;
    double bcur[NB];
    const int nchunk = (np + KC - 1) / KC;
    if (wave_active) {
#pragma unroll
        for (int k = 0; k < NB; ++k) bcur[k] = cbw[(4 * k) * 16];
    }
    // panel buffers: piece p, chunk parity q at (2 p + q) * PANEL
    fetch(0, zsrc0, zst0);
    if (two) fetch(0, zsrc1, zst1);
    commit(0, 0, zst0);
    if (two) commit(0, 2 * PANEL, zst1);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
    for (int ch = 0; ch < nchunk; ++ch) {
        const bool more = (ch + 1) < nchunk;
        double bnext[NB];
        if (more && wave_active) {
#pragma unroll
            for (int k = 0; k < NB; ++k) bnext[k] = cbw[((ch + 1) * KC + 4 * k) * 16];
        }
        __builtin_amdgcn_sched_barrier(0);
        if (more) {
            fetch((ch + 1) * KC, zsrc0, zst0);
            if (two) fetch((ch + 1) * KC, zsrc1, zst1);
        }
        __builtin_amdgcn_sched_barrier(0);
        if (wave_active) {
            const int zo = (2 * piece + (ch & 1)) * PANEL;
#pragma unroll
            for (int k = 0; k < NB; ++k) {
#pragma unroll
                for (int rb = 0; rb < NRB; ++rb) {
                    const double af = smem[zo + (4 * k + lg) * LDZ + rb * 16 + lm];
                    acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, bcur[k], acc[rb], 0, 0, 0);
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        if (more) {
            commit((ch + 1) * KC, ((ch + 1) & 1) * PANEL, zst0);
            if (two) commit((ch + 1) * KC, (2 + ((ch + 1) & 1)) * PANEL, zst1);
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        if (more && wave_active) {
#pragma unroll
            for (int k = 0; k < NB; ++k) bcur[k] = bnext[k];
        }
    }
    if (!wave_active) return;

    // Q4 as in half_tile: O^T = T^T C, a <= c, pair0(c, a) < limit
    const int my_last_ab = cw / 16;
    gdouble* __restrict__ dst = (gdouble*)A.dst + dst_off;
    constexpr int QD = 3;
    double ring[QD + 1][4];
    gcdouble* __restrict__ cq = cr + (int64_t)kCRows * A.ldm;
    auto load_rb = [&](int ab, int rb, double* bq) {
        gcdouble2* __restrict__ p = (gcdouble2*)(cq + ((int64_t)ab * 16 + rb) * 256) + lane;
        const f64x2 v0 = p[0], v1 = p[64];
        bq[0] = v0.x;
        bq[1] = v0.y;
        bq[2] = v1.x;
        bq[3] = v1.y;
    };
#pragma unroll
    for (int rb = 0; rb < QD && rb < NRB; ++rb) load_rb(0, rb, ring[rb]);
    for (int ab = 0; ab <= my_last_ab; ++ab) {
        f64x4 o0 = f64x4{0.0, 0.0, 0.0, 0.0}, o1 = o0;
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb) {
            if (rb + QD < NRB) load_rb(ab, rb + QD, ring[(rb + QD) % (QD + 1)]);
            const double* bq = ring[rb % (QD + 1)];
            o0 = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[rb][0], bq[0], o0, 0, 0, 0);
            o1 = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[rb][1], bq[1], o1, 0, 0, 0);
            o0 = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[rb][2], bq[2], o0, 0, 0, 0);
            o1 = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[rb][3], bq[3], o1, 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (ab < my_last_ab) {
#pragma unroll
            for (int rb = 0; rb < QD && rb < NRB; ++rb) load_rb(ab + 1, rb, ring[rb]);
        }
        __builtin_amdgcn_sched_barrier(0);
        const int a = ab * 16 + lm;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = cw + lg + 4 * q;
            if (c < c1 && a <= c) {
                const int64_t pp = pair0(c, a);
                if (pp < limit) dst[pp] = o0[q] + o1[q];
            }
        }
    }
}

template <int NRB>
__global__ __launch_bounds__(512, 1) void k_half2_pair(HalfArgs A, const PairItem* __restrict__ items,
                                                      int64_t nitems) {
    const int64_t item = xcd_item(blockIdx.x);
    if (item >= nitems) return;
    const PairItem it = items[item];
    pair_tile<NRB>(A, it);
}

template <int NRB>
static hipError_t launch_pair_rows(const HalfArgs& a, const PairItem* items, int64_t nitems,
                                   hipStream_t st) {
    const size_t shmem = sizeof(double) * 4 * 16 * (NRB * 16 + 1);
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)k_half2_pair<NRB>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    const int64_t g = xcd_grid(nitems);
    if (g == 0) return hipSuccess;
    if (g > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_half2_pair<NRB>), dim3((unsigned)g), dim3(512), shmem, st, a, items, nitems);
    return hipGetLastError();
}

// Paired half 2 grouped over the transforms of one row class: an item names its transform
// (an index into geoms) besides its pieces, and the pieces of one item come from one
// transform.
template <int NRB>
__global__ __launch_bounds__(512, 1) void k_half2_pair_group(const HalfArgs* __restrict__ geoms,
                                                            const GroupPairItem* __restrict__ items,
                                                            int64_t nitems) {
    const int64_t item = xcd_item(blockIdx.x);
    if (item >= nitems) return;
    const GroupPairItem gi = items[item];
    const HalfArgs A = geoms[gi.geom];
    pair_tile<NRB>(A, gi.p);
}

template <int NRB>
static hipError_t launch_group_pair_rows(const GroupLaunch& gl, hipStream_t st) {
    const void* fn = (const void*)k_half2_pair_group<NRB>;
    const size_t shmem = sizeof(double) * 4 * 16 * (NRB * 16 + 1);
    if (gl.grid <= 0)
        return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (!gl.pair_items || gl.grid > 0x7fffffff || gl.grid != xcd_grid(gl.grid) ||
        gl.grid < gl.npair)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_half2_pair_group<NRB>), dim3((unsigned)gl.grid), dim3(512), shmem, st,
                       gl.geoms, gl.pair_items, gl.npair);
    return hipGetLastError();
}


template <int NW, int NRB, int KC, int MINB>
__global__ __launch_bounds__(NW * 64, MINB) void k_half_fp64(HalfArgs A) {
    const int64_t item = A.item_begin + xcd_item(blockIdx.x);
    if (item >= A.ntile_total) return;
    half_tile<NW, NRB, KC>(A, item);
}

// Grouped launch over the work items of several transforms (a scan of geometries).
// Items are numbered geometry after geometry (item_start[g] is the first item of
// geometry g, item_start[ngeom] the total); the grid holds one work-group per item, so
// the hardware dispatcher balances the work as in a single transform's launch, with the
// same chunked XCD mapping.  The owning geometry is found by binary search over the
// start table.
template <int NW, int NRB, int KC, int MINB>
__global__ __launch_bounds__(NW * 64, MINB) void k_half_fp64_group(const HalfArgs* __restrict__ geoms,
                                                                  const int64_t* __restrict__ item_start,
                                                                  int ngeom) {
    const int64_t total = item_start[ngeom];
    const int64_t it = xcd_item(blockIdx.x);
    if (it >= total) return;
    int lo = 0, hi = ngeom;                        // item_start[0] = 0 <= it < total
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (item_start[mid] <= it) lo = mid; else hi = mid;
    }
    const HalfArgs A = geoms[lo];
    half_tile<NW, NRB, KC>(A, it - item_start[lo]);
}

template <int NW, int NRB, int KC>
static size_t half_shmem() {
    return sizeof(double) * 2 * (size_t)KC * (NRB * 16 + 1);
}

template <int NW, int NRB, int KC, int MINB>
static hipError_t launch_v(const HalfArgs& a, hipStream_t st) {
    using K = HalfCfg<NW, NRB, KC>;
    const size_t shmem = half_shmem<NW, NRB, KC>();
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)k_half_fp64<NW, NRB, KC, MINB>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    if (a.item_begin < 0 || a.item_begin > a.ntile_total) return hipErrorInvalidValue;
    const int64_t g = xcd_grid(a.ntile_total - a.item_begin);
    if (g == 0) return hipSuccess;
    if (g > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_half_fp64<NW, NRB, KC, MINB>), dim3((unsigned)g), dim3(K::NT), shmem, st, a);
    return hipGetLastError();
}

// Grouped launch.  grid <= 0 only prepares the kernel (LDS attribute); grid > 0
// launches (one work-group per item, rounded up by xcd_grid).
template <int NW, int NRB, int KC, int MINB>
static hipError_t launch_group_v(const GroupLaunch& gl, hipStream_t st) {
    using K = HalfCfg<NW, NRB, KC>;
    const void* fn = (const void*)k_half_fp64_group<NW, NRB, KC, MINB>;
    const size_t shmem = half_shmem<NW, NRB, KC>();
    if (gl.grid <= 0)
        return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (gl.grid > 0x7fffffff || gl.grid != xcd_grid(gl.grid)) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_half_fp64_group<NW, NRB, KC, MINB>), dim3((unsigned)gl.grid), dim3(K::NT),
                       shmem, st, gl.geoms, gl.item_start, gl.ngeom);
    return hipGetLastError();
}

// Tile variants: (waves per work-group, Q1 panel depth, work-groups per CU hint).
// Columns per work-group = 16 x waves.  The 12-wave variant runs three waves per SIMD,
// which caps a wave at 170 registers: its T accumulators fit for up to 12 row blocks
// (N_AO <= 192); with more it would spill, so it is not built for those row classes.
const HalfVariant kHalfVariants[] = {
    {4, 16, 1}, {4, 32, 1}, {8, 16, 1}, {8, 32, 1}, {2, 16, 1}, {4, 16, 2}, {8, 8, 1},
    {12, 16, 1},
};
const int kNumHalfVariants = sizeof(kHalfVariants) / sizeof(kHalfVariants[0]);
constexpr int kMaxRows12 = 12;

bool variant_supports(int v, int row_cls) {
    if (v < 0 || v >= kNumHalfVariants) return false;
    return kHalfVariants[v].waves != 12 || row_cls <= kMaxRows12;
}

template <int NRB>
static hipError_t launch_rows(const HalfArgs& a, int v, hipStream_t st) {
    switch (v) {
        case 0: return launch_v<4, NRB, 16, 1>(a, st);
        case 1: return launch_v<4, NRB, 32, 1>(a, st);
        case 2: return launch_v<8, NRB, 16, 1>(a, st);
        case 3: return launch_v<8, NRB, 32, 1>(a, st);
        case 4: return launch_v<2, NRB, 16, 1>(a, st);
        case 5: return launch_v<4, NRB, 16, 2>(a, st);
        case 6: return launch_v<8, NRB, 8, 1>(a, st);
        case 7:
            if constexpr (NRB <= kMaxRows12) return launch_v<12, NRB, 16, 1>(a, st);
            else return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

template <int NRB>
static hipError_t launch_group_rows(const GroupLaunch& gl, hipStream_t st) {
    switch (gl.variant) {
        case 0: return launch_group_v<4, NRB, 16, 1>(gl, st);
        case 1: return launch_group_v<4, NRB, 32, 1>(gl, st);
        case 2: return launch_group_v<8, NRB, 16, 1>(gl, st);
        case 3: return launch_group_v<8, NRB, 32, 1>(gl, st);
        case 4: return launch_group_v<2, NRB, 16, 1>(gl, st);
        case 5: return launch_group_v<4, NRB, 16, 2>(gl, st);
        case 6: return launch_group_v<8, NRB, 8, 1>(gl, st);
        case 7:
            if constexpr (NRB <= kMaxRows12) return launch_group_v<12, NRB, 16, 1>(gl, st);
            else return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

// Compiled row-block counts: every count from 4 to 16 (N_AO up to 256), so a transform
// computes no padded rows beyond its last 16-row block.
int row_class(int np) {
    const int nrb = np / 16;
    return nrb < 4 ? 4 : nrb;
}

#define CIPSI_ROW_CLASSES(X) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)

hipError_t launch_half_fp64(const HalfArgs& a, int variant, hipStream_t st) {
    if (a.np > 256 || (a.np & 15) || (a.mp & 15)) return hipErrorInvalidValue;
    switch (row_class(a.np)) {
#define CIPSI_CASE(R) case R: return launch_rows<R>(a, variant, st);
        CIPSI_ROW_CLASSES(CIPSI_CASE)
#undef CIPSI_CASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_half2_pair(const HalfArgs& a, const PairItem* items, int64_t nitems,
                             hipStream_t st) {
    if (a.half != 2 || a.np > 256 || (a.np & 15) || (a.mp & 15)) return hipErrorInvalidValue;
    switch (row_class(a.np)) {
#define CIPSI_CASE(R) case R: return launch_pair_rows<R>(a, items, nitems, st);
        CIPSI_ROW_CLASSES(CIPSI_CASE)
#undef CIPSI_CASE
        default: return hipErrorInvalidValue;
    }
}

// Work items of the paired half 2 (see pair_tile) for the MO pair rows l0 <= l < l1,
// appended to out in the order of the rows: full pieces as their matrix comes, a partial
// piece or remainder when a waiting piece completes it to 8 blocks, and the pieces still
// waiting at the end paired largest first or alone.  Within an item the waves take the
// blocks as in half_tile: sorted by block, SIMD j < P = nact - 4 holds the j-th smallest
// and the j-th largest, the other SIMDs one block each.
void build_pair_items(int l0, int l1, std::vector<PairItem>& out) {
    struct Piece { int64_t b; int first, count; };   // blocks [first, first + count) of b
    std::vector<Piece> waiting[8];
    auto emit = [&](const Piece* p, int np_) {
        // entries (block, piece), sorted by block
        int blk[8], pc[8], ne = 0;
        for (int i = 0; i < np_; ++i)
            for (int k = 0; k < p[i].count; ++k) { blk[ne] = p[i].first + k; pc[ne] = i; ++ne; }
        for (int i = 1; i < ne; ++i)
            for (int j = i; j > 0 && blk[j] < blk[j - 1]; --j) {
                std::swap(blk[j], blk[j - 1]);
                std::swap(pc[j], pc[j - 1]);
            }
        PairItem it;
        it.b[0] = p[0].b;
        it.b[1] = np_ > 1 ? p[1].b : -1;
        it.slots = 0;
        const int p2 = ne > 4 ? ne - 4 : 0;
        for (int wv = 0; wv < 8; ++wv) {
            const int j = wv & 3;
            int e;
            if (j < p2) e = (wv >> 2) ? 2 * p2 - 1 - j : j;
            else e = (wv >> 2) ? -1 : p2 + j;
            if (e < 0 || e >= ne) continue;
            const uint64_t code = 0x80u | ((unsigned)pc[e] << 4) | (unsigned)blk[e];
            it.slots |= code << (8 * wv);
        }
        out.push_back(it);
    };
    auto add_partial = [&](Piece pc) {
        const int s = pc.count;
        if (s == 8) {
            emit(&pc, 1);
        } else if (!waiting[8 - s].empty()) {
            Piece two[2] = {waiting[8 - s].back(), pc};
            waiting[8 - s].pop_back();
            emit(two, 2);
        } else {
            waiting[s].push_back(pc);
        }
    };
    for (int l = l0; l < l1; ++l) {
        const int nb = (l + 16) / 16;
        for (int k = 0; k <= l; ++k) {
            const int64_t b = pair0(l, k);
            if (nb == 8) {
                Piece p{b, 0, 8};
                emit(&p, 1);
            } else if (nb < 8) {
                add_partial(Piece{b, 0, nb});
            } else {
                // full piece: blocks 0..3 and 4 + r .. nb - 1; remainder: blocks 4 .. 3 + r
                const int r = nb - 8;
                PairItem it;
                it.b[0] = b;
                it.b[1] = -1;
                it.slots = 0;
                int blk[8];
                for (int i = 0; i < 4; ++i) blk[i] = i;
                for (int i = 0; i < 4; ++i) blk[4 + i] = 4 + r + i;
                for (int j = 0; j < 4; ++j) {           // SIMD j: the j-th smallest and largest
                    it.slots |= (uint64_t)(0x80u | (unsigned)blk[j]) << (8 * j);
                    it.slots |= (uint64_t)(0x80u | (unsigned)blk[7 - j]) << (8 * (j + 4));
                }
                out.push_back(it);
                add_partial(Piece{b, 4, r});
            }
        }
    }
    // leftovers: pair any two that fit, largest first, then singles
    for (int s = 7; s >= 1; --s)
        while (!waiting[s].empty()) {
            Piece a0 = waiting[s].back();
            waiting[s].pop_back();
            int t = 8 - s;
            while (t >= 1 && waiting[t].empty()) --t;
            if (t >= 1) {
                Piece two[2] = {a0, waiting[t].back()};
                waiting[t].pop_back();
                emit(two, 2);
            } else {
                emit(&a0, 1);
            }
        }
}

hipError_t launch_half_fp64_group(const GroupLaunch& gl, hipStream_t st) {
    if (gl.paired) {
        if (gl.variant != kPairedHalf2Variant) return hipErrorInvalidValue;
        switch (gl.row_class) {
#define CIPSI_CASE(R) case R: return launch_group_pair_rows<R>(gl, st);
            CIPSI_ROW_CLASSES(CIPSI_CASE)
#undef CIPSI_CASE
            default: return hipErrorInvalidValue;
        }
    }
    switch (gl.row_class) {
#define CIPSI_CASE(R) case R: return launch_group_rows<R>(gl, st);
        CIPSI_ROW_CLASSES(CIPSI_CASE)
#undef CIPSI_CASE
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cipsi
