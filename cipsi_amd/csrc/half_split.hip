// Split-FP16 (Ozaki) half-transform on FP16 MFMA with FP32 accumulation.
//
// Each half (tpqkld.f:82-97 / tijkld.f:53-76) is run as two slice products:
//   step 1 (Q1/Q3)  T_b = Z_b C       Z_b rows gathered from the packed symmetric rows;
//                                     stored transposed, U[b][c][r] = T_b(r, c)
//   step 2 (Q2/Q4)  O_b^T = U_b C     rows (b, c) of U; only a <= c is kept and written
//                                     to the packed output exactly as the FP64 kernel does.
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
//   out = sum_d 2^(e_row + f_col + 2 - 8d) acc_d  in FP64.
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
constexpr int TR = 32;        // rows of A per work-group
constexpr int KCH = 32;       // K per chunk (one MFMA)
constexpr int APAD = 40;      // LDS row stride of a digit plane, in halves
constexpr int NWAVE = 8;      // 8 waves x 32 columns = up to 256 output columns
constexpr int NTHR = NWAVE * 64;

typedef __attribute__((address_space(1))) const double gcdouble;

__host__ __device__ inline int round32(int x) { return (x + 31) & ~31; }

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
// work-group per matrix, coalesced over the packed row, LDS atomic max per AO index.
__global__ __launch_bounds__(256) void k_split_rowexp(const double* __restrict__ src,
                                                      int64_t src_ld, int n, int kp,
                                                      int* __restrict__ rexp) {
    __shared__ int emax[256];
    const int64_t b = blockIdx.x;
    for (int i = threadIdx.x; i < 256; i += 256) emax[i] = INT_MIN;
    __syncthreads();
    const int64_t pa = npairs(n);
    const double* row = src + b * src_ld;
    for (int64_t idx = threadIdx.x; idx < pa; idx += 256) {
        const double x = row[idx];
        if (x == 0.0) continue;
        const int r = (int)pair_row(idx);
        const int s = (int)(idx - pair0(r, 0));
        const int e = exp_above(x);
        atomicMax(&emax[r], e);
        if (s != r) atomicMax(&emax[s], e);
    }
    __syncthreads();
    for (int r = threadIdx.x; r < kp; r += 256) {
        const int e = r < n ? emax[r] : INT_MIN;
        rexp[b * kp + r] = e == INT_MIN ? 0 : e;
    }
}

__global__ void k_fill_int(int* __restrict__ p, int64_t n, int v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

struct SplitArgs {
    const double* src;       // step 1: packed symmetric rows; step 2: U rows (length kp)
    int64_t src_ld;
    const int* rexp;         // step 1: [b][kp] row exponents; step 2: [b][mp] U row exponents
    const _Float16* cs;      // C digits
    const int* cexp;         // C column exponents
    double* dst;             // step 1: U [b][mp][kp]; step 2: packed output
    int* dexp;               // step 1: U row exponents [b][mp] (atomic max)
    int n, m, mp, kp;
    int half;
    int64_t nb;
    int64_t dst_ld;          // step 2, half 1: P_M
    const int64_t* tile_start;  // step 2, half 2: prefix over l of (l+1) ceil((l+1)/32)
    int64_t nitems;
};

// One work item: 32 rows of A (matrix b) times up to 256 columns of C.
template <int S, int STEP>
__global__ __launch_bounds__(NTHR) void k_split_gemm(SplitArgs A) {
    __shared__ _Float16 As[2][S][TR][APAD];
    __shared__ int rx[TR];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int lm = lane & 15, lg = lane >> 4;
    const int64_t item = xcd_item(blockIdx.x);
    if (item >= A.nitems) return;

    // ---- work item -> matrix b, row tile, extents
    int64_t b;
    int rt, nrows, ncols;
    int64_t limit = 0, dst_off = 0;
    const int kc = A.kp / KCH;
    if (STEP == 1) {
        const int ntr = A.kp / TR;
        b = item / ntr;
        rt = (int)(item % ntr);
        nrows = A.n;
        ncols = A.half == 1 ? A.m : (int)pair_row(b) + 1;        // half 2: columns j <= l
    } else {
        if (A.half == 1) {
            const int ntr = (A.m + TR - 1) / TR;
            b = item / ntr;
            rt = (int)(item % ntr);
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
            const int nt = (l + TR) / TR;
            const int64_t rem = item - A.tile_start[l];
            b = pair0(l, 0) + rem / nt;
            rt = (int)(rem % nt);
            nrows = l + 1;
            limit = b + 1;
            dst_off = pair0(b, 0);
        }
        ncols = min(rt * TR + TR, nrows);                           // a <= c
    }
    const int row0 = rt * TR;
    const int cw = 32 * w;                                          // this wave's columns
    const bool active = cw < ncols;

    if (tid < TR) {
        const int r = row0 + tid;
        int e = 0;
        if (STEP == 1) e = r < A.n ? A.rexp[b * A.kp + r] : 0;
        else if (r < nrows) {
            e = A.rexp[b * A.mp + r];
            if (e == INT_MIN) e = 0;
        }
        rx[tid] = e;
    }
    __syncthreads();

    // ---- A chunk staging: element u of this thread -> (row i, k)
    constexpr int EPT = TR * KCH / NTHR;                            // 2
    double xs[EPT];
    gcdouble* __restrict__ src = (gcdouble*)A.src;
    // Loads are unconditional (an element outside the matrix reads element 0 of its row,
    // which exists) and masked at commit, so no register is written by anything but its
    // load and the loop body has no branches around loads.
    auto elem = [&](int u, int ch, int64_t& idx) -> bool {
        const int e = tid + u * NTHR;
        const int i = e / KCH, k = e % KCH;
        const int r = row0 + i, s = ch * KCH + k;
        bool ok;
        if (STEP == 1) {
            ok = r < A.n && s < A.n;
            const int hi = r >= s ? r : s, lo = r >= s ? s : r;
            idx = b * A.src_ld + (ok ? pair0(hi, lo) : 0);
        } else {
            ok = r < nrows && s < A.n;
            idx = ok ? (b * A.mp + r) * (int64_t)A.kp + s : b * A.mp * (int64_t)A.kp;
        }
        return ok;
    };
    auto fetch = [&](int ch) {
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            int64_t idx;
            (void)elem(u, ch, idx);
            xs[u] = src[idx];
        }
    };
    auto commit = [&](int ch, int buf) {
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            const int e = tid + u * NTHR;
            const int i = e / KCH, k = e % KCH;
            int64_t idx;
            const bool ok = elem(u, ch, idx);
            _Float16 d[S];
            digits<S>(ok ? xs[u] : 0.0, rx[i], d);
#pragma unroll
            for (int s = 0; s < S; ++s) As[buf][s][i][k] = d[s];
        }
    };

    f32x4 acc[S][2][2];
#pragma unroll
    for (int g = 0; g < S; ++g)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[g][i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    typedef __attribute__((address_space(1))) const f16x8 gcf16x8;
    gcf16x8* __restrict__ csv = (gcf16x8*)A.cs;

    // B fragments (digits of C), double-buffered in registers: the next chunk's S x 2
    // fragments are loaded at the start of a chunk, ahead of its A prefetch, by every
    // wave (a wave without columns reads a clamped column), so that no load sits in a
    // branch.  Loads complete in order: the commit waits only for the A prefetch, and a
    // chunk's MFMAs wait only for B loads issued a whole chunk earlier.  Columns past mp
    // only feed outputs that are never stored.  The A fragments are read from LDS per
    // digit i, so only one pair is live.
    f16x8 bcur[S][2], bnext[S][2];
    auto load_b = [&](int ch, f16x8 (&bf)[S][2]) {
#pragma unroll
        for (int j = 0; j < S; ++j)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                const int c = min(cw + 16 * cb + lm, A.mp - 1);
                bf[j][cb] = csv[(((int64_t)j * kc + ch) * A.mp + c) * (KCH / 8) + lg];
            }
    };
    auto mfma_chunk = [&](int ch) {
        const int buf = ch & 1;
#pragma unroll
        for (int i = 0; i < S; ++i) {
            f16x8 af[2];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
                af[rb] = *(const f16x8*)&As[buf][i][16 * rb + lm][8 * lg];
#pragma unroll
            for (int j = 0; i + j < S; ++j)
#pragma unroll
                for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                    for (int cb = 0; cb < 2; ++cb)
                        acc[i + j][rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                            af[rb], bcur[j][cb], acc[i + j][rb][cb], 0, 0, 0);
        }
    };
    load_b(0, bcur);
    fetch(0);
    commit(0, 0);
    __builtin_amdgcn_s_waitcnt(0x0F70);            // vmcnt(0): nothing pending into the loop
    __syncthreads();
    // all chunks but the last prefetch the next chunk's B and A; the last chunk is peeled
    // so that the loop body has no conditional loads
    for (int ch = 0; ch + 1 < kc; ++ch) {
        load_b(ch + 1, bnext);
        fetch(ch + 1);
        __builtin_amdgcn_sched_barrier(0);
        if (active) mfma_chunk(ch);
        __builtin_amdgcn_sched_barrier(0);
        commit(ch + 1, (ch + 1) & 1);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < S; ++j)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) bcur[j][cb] = bnext[j][cb];
    }
    if (active) mfma_chunk(kc - 1);
    if (!active) return;

    // ---- epilogue: FP64 recombination and stores
    typedef __attribute__((address_space(1))) double gdouble;
    gdouble* __restrict__ dst = (gdouble*)A.dst;
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        const int col = cw + 16 * cb + lm;
        const int fc = col < A.mp ? A.cexp[col] : 0;
        int emax = INT_MIN;
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int il = 16 * rb + 4 * lg + q;
                const int base = rx[il] + fc + 2;
                double v = 0.0;
#pragma unroll
                for (int g = S - 1; g >= 0; --g)
                    v += ldexp((double)acc[g][rb][cb][q], base - 8 * (g + 2));
                const int r = row0 + il;
                if (STEP == 1) {
                    if (r < A.n && col < ncols) {
                        dst[(b * A.mp + col) * (int64_t)A.kp + r] = v;
                        if (v != 0.0) emax = max(emax, exp_above(v));
                    }
                } else {
                    if (r < nrows && col <= r) {
                        const int64_t p = pair0(r, col);
                        if (p < limit) dst[dst_off + p] = v;
                    }
                }
            }
        }
        if (STEP == 1) {
            // row exponent of U row (b, col): max over the 32 rows of this tile
            emax = max(emax, __shfl_xor(emax, 16));
            emax = max(emax, __shfl_xor(emax, 32));
            if (lg == 0 && col < ncols && emax != INT_MIN) atomicMax(&A.dexp[b * A.mp + col], emax);
        }
    }
}

template <int S, int STEP>
hipError_t launch_split_gemm(const SplitArgs& a, hipStream_t st) {
    const int64_t g = xcd_grid(a.nitems);
    if (g <= 0 || g > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_split_gemm<S, STEP>), dim3((unsigned)g), dim3(NTHR), 0, st, a);
    return hipGetLastError();
}

template <int STEP>
hipError_t launch_split_s(int s, const SplitArgs& a, hipStream_t st) {
    switch (s) {
        case 3: return launch_split_gemm<3, STEP>(a, st);
        case 4: return launch_split_gemm<4, STEP>(a, st);
        case 5: return launch_split_gemm<5, STEP>(a, st);
        case 6: return launch_split_gemm<6, STEP>(a, st);
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
    const int kp = round32(n), mp = (m + 15) & ~15;
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

int run_half_split(cipsi_ijkl_ctx* ctx, const HalfArgs& a, int, int n_slices) {
    const int s = n_slices == 0 ? kDefaultSlices : n_slices;
    if (s < kMinSlices || s > kMaxSlices) return CIPSI_ERR_UNSUPPORTED;
    const int n = a.n, m = a.m, mp = a.mp, kp = round32(n);
    if (mp > NWAVE * 32 || kp > 256) return CIPSI_ERR_UNSUPPORTED;
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
    hipLaunchKernelGGL(k_fill_int, dim3(1024), dim3(256), 0, st, uexp, a.nb * (int64_t)mp, INT_MIN);
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
    x.nitems = a.nb * (int64_t)(kp / TR);
    if ((e = launch_split_s<1>(s, x, st)) != hipSuccess) return CIPSI_ERR_HIP;
    // step 2: O_b^T = U_b C, packed triangle
    x.src = u;
    x.src_ld = 0;
    x.rexp = uexp;
    x.dst = a.dst;
    x.dexp = nullptr;
    if (a.half == 1) {
        x.dst_ld = a.dst_ld;
        x.tile_start = nullptr;
        x.nitems = a.nb * (int64_t)((m + TR - 1) / TR);
    } else {
        const int64_t* ts;
        int64_t total;
        if ((rc = ctx_tile_table(ctx, m, TR, &ts, &total))) return rc;
        x.tile_start = ts;
        x.nitems = total;
    }
    if ((e = launch_split_s<2>(s, x, st)) != hipSuccess) return CIPSI_ERR_HIP;
    return CIPSI_OK;
}
