// HBM-bound layout kernels of the ijkl transform (gfx950).
//
//  unpack_labels  : pshf label stream -> packed AO block            (transd.f, C1 route)
//  expand_packed  : packed lower-triangular AO block -> full symmetric pair matrix
//  transpose      : H[pq][kl] -> H^T[kl][pq] between the two halves
//  pad_c          : C (column-major, ldc) -> zero-padded copy in 16-column blocks of kCRows
//                   rows (cr_index) read by the MFMA tiles, and its second-product order
//                   (cq_index)
#include "common.h"

namespace cipsi {

// ----------------------------------------------------------------------------- fill
__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

static inline unsigned grid_for(int64_t n, int threads, unsigned cap = 65536u) {
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    return (unsigned)(g > cap ? cap : g);
}


// ----------------------------------------------------------------------------- labels
// transd.f:85-97 (decode, first zero label ends the stream), 210-218 (doubling of
// the i==j, k==l and ij==kl entries that pshf/qout.f:73-77 halved), 369-372 (store
// at ij1(ij1-1)/2 + kl1), into a block zeroed first (itijkl.f:191-192).  The reference
// processes labels sequentially, so a later duplicate overwrites an earlier one: the
// last writer of each position is elected with atomicMax on (index + 1).
//
// Passes (all streaming in the order pshf writes the labels):
//   claim   every label of the buffer: a zero label lowers the stream end, a label with
//           a field outside 1..n_ao lowers the first bad index, a valid one claims its
//           position and raises the last valid index;
//   fix     only when a valid label lies after the first zero (the stale tail of the
//           last record): the owners are cleared and re-claimed up to the stream end;
//   resolve one thread per position: the owner's value times 2^(doubling), or 0.
//   noncanon only when a valid label is not in pshf's canonical order (i >= j, k >= l,
//           ij >= kl): resolve derives the doubling from the position, which for such a
//           label differs from transd's, which tests the label's own fields
//           (transd.f:211-218); the owners among them are rewritten with transd's value.
// Scalars sc[]: 0 = stream end (first zero label), 1 = first bad label, 2 = last valid
// label + 1 (0 = none), 3 = nonzero if any valid label is non-canonical, all as unsigned
// 64-bit for the atomics.

__global__ void k_init_label_scalars(unsigned long long* sc, int64_t n, int* err_flag) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sc[0] = (unsigned long long)n;
        sc[1] = (unsigned long long)n;
        sc[2] = 0ull;
        sc[3] = 0ull;
        *err_flag = 0;
    }
}

// Fields of a label (transd.f:94-97), 1-based.
__device__ inline void label_fields(int32_t lab, int& i, int& j, int& k, int& l) {
    const uint32_t u = (uint32_t)lab;
    i = (int)((u >> 24) & 255u);
    j = (int)((u >> 16) & 255u);
    k = (int)((u >> 8) & 255u);
    l = (int)(u & 255u);
}

__device__ inline bool label_canonical(int i, int j, int k, int l) {
    return i >= j && k >= l && i * (i - 1) / 2 + j >= k * (k - 1) / 2 + l;
}

// 1-based fields (transd.f:91-97) -> 0-based packed position; false if out of range.
__device__ inline bool decode_label(int32_t lab, int n_ao, int64_t* pos) {
    int i, j, k, l;
    label_fields(lab, i, j, k, l);
    if (i < 1 || j < 1 || k < 1 || l < 1 || i > n_ao || j > n_ao || k > n_ao || l > n_ao)
        return false;
    const int64_t ij1 = (int64_t)i * (i - 1) / 2 + j;
    const int64_t kl1 = (int64_t)k * (k - 1) / 2 + l;
    const int64_t p = ij1 * (ij1 - 1) / 2 + kl1 - 1;   // 0-based
    if (p < 0 || p >= npairs(npairs(n_ao))) return false;
    *pos = p;
    return true;
}

constexpr int kLabelThreads = 256;

// Block-wide min / max of per-thread values, then one atomic per block.
__device__ inline void block_commit(unsigned long long zmin, unsigned long long bmin,
                                    unsigned long long vmax, unsigned long long* sc) {
    __shared__ unsigned long long r[3][kLabelThreads / 64];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        zmin = min(zmin, (unsigned long long)__shfl_xor(zmin, o));
        bmin = min(bmin, (unsigned long long)__shfl_xor(bmin, o));
        vmax = max(vmax, (unsigned long long)__shfl_xor(vmax, o));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { r[0][w] = zmin; r[1][w] = bmin; r[2][w] = vmax; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kLabelThreads / 64; ++i) {
            zmin = min(zmin, r[0][i]);
            bmin = min(bmin, r[1][i]);
            vmax = max(vmax, r[2][i]);
        }
        if (zmin < sc[0]) atomicMin(&sc[0], zmin);
        if (bmin < sc[1]) atomicMin(&sc[1], bmin);
        if (vmax > sc[2]) atomicMax(&sc[2], vmax);
    }
}

// claim over every label of the buffer.  Consecutive lanes take consecutive labels, so
// the atomics of a wave instruction fall on as few cache lines as the stream allows.
__global__ __launch_bounds__(kLabelThreads) void k_claim_scan(const int32_t* __restrict__ labels,
                                                               int64_t n, int n_ao,
                                                               uint32_t* __restrict__ owner,
                                                               unsigned long long* sc) {
    unsigned long long zmin = ~0ull, bmin = ~0ull, vmax = 0ull;
    bool nonc = false;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int32_t lab = labels[i];
        int64_t pos;
        if (lab == 0) {
            zmin = min(zmin, (unsigned long long)i);
        } else if (!decode_label(lab, n_ao, &pos)) {
            bmin = min(bmin, (unsigned long long)i);
        } else {
            atomicMax(owner + pos, (uint32_t)(i + 1));
            vmax = max(vmax, (unsigned long long)(i + 1));
            int a, b, c, d;
            label_fields(lab, a, b, c, d);
            nonc = nonc || !label_canonical(a, b, c, d);
        }
    }
    if (__ballot(nonc) != 0ull && (threadIdx.x & 63) == 0) atomicOr(&sc[3], 1ull);
    block_commit(zmin, bmin, vmax, sc);
}

// Non-canonical owners: the value transd stores, xx doubled for i == j, k == l and
// (i, j) == (k, l) of the label itself (transd.f:210-218), over the resolve's value.
__global__ void k_noncanon_fix(const double* __restrict__ xx, const int32_t* __restrict__ labels,
                               int n_ao, const uint32_t* __restrict__ owner,
                               const unsigned long long* sc, double* __restrict__ out) {
    if (sc[3] == 0ull) return;
    const int64_t n = (int64_t)sc[0];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        const int32_t lab = labels[t];
        int64_t pos;
        if (!decode_label(lab, n_ao, &pos)) continue;
        int i, j, k, l;
        label_fields(lab, i, j, k, l);
        if (label_canonical(i, j, k, l) || owner[pos] != (uint32_t)(t + 1)) continue;
        double x = xx[t];
        if (i == j) x = x + x;
        if (k == l) x = x + x;
        if (i == k && j == l) x = x + x;
        out[pos] = x;
    }
}

// A valid label after the stream end claimed a position: clear and re-claim up to the end.
__global__ void k_fix_clear(uint32_t* __restrict__ owner, int64_t nblk,
                            const unsigned long long* sc) {
    if (sc[2] <= sc[0] + 1) return;                  // last valid index < stream end
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nblk; p += stride)
        owner[p] = 0u;
}

__global__ void k_fix_claim(const int32_t* __restrict__ labels, int n_ao,
                            uint32_t* __restrict__ owner, const unsigned long long* sc) {
    if (sc[2] <= sc[0] + 1) return;
    const int64_t n = (int64_t)sc[0];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        int64_t pos;
        if (decode_label(labels[i], n_ao, &pos)) atomicMax(owner + pos, (uint32_t)(i + 1));
    }
}

// Largest r with r(r+1)/2 <= x, from a single-precision estimate made exact with
// integer steps (x < 2^31 here: the packed block of N_AO <= 255 has 5.3e8 entries).
__device__ inline int64_t tri_row(int64_t x) {
    int64_t r = (int64_t)((sqrtf(8.0f * (float)x + 1.0f) - 1.0f) * 0.5f);
    while (r * (r + 1) / 2 > x) --r;
    while ((r + 1) * (r + 2) / 2 <= x) ++r;
    return r;
}

// One thread per four consecutive positions: the first is decomposed into
// (ab, cd) = pair indices and (i, j), (k, l) = AO indices, the next three by stepping.
// The doubling count is (i == j) + (k == l) + (ab == cd) (transd.f:210-218).
// The owner quad of the thread's next four positions is loaded one iteration ahead, so
// an owner load is in flight while the value gathers of the current quad wait.  With a
// 16-byte-aligned output (VEC) a quad is stored as two 16-byte writes.
__device__ __forceinline__ void load_owner4(const uint32_t* __restrict__ owner, int64_t p0,
                                            int64_t nblk, uint32_t w[4]) {
    if (p0 + 3 < nblk) {
        const uint4 o = *(const uint4*)(owner + p0);             // owner is 16-byte aligned
        w[0] = o.x; w[1] = o.y; w[2] = o.z; w[3] = o.w;
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) w[e] = p0 + e < nblk ? owner[p0 + e] : 0u;
    }
}

template <bool VEC>
__global__ __launch_bounds__(kLabelThreads) void k_resolve(const double* __restrict__ xx,
                                                            int64_t nblk,
                                                            const uint32_t* __restrict__ owner,
                                                            const unsigned long long* sc,
                                                            int* err_flag,
                                                            double* __restrict__ out) {
    const uint64_t n = sc[0];
    if (blockIdx.x == 0 && threadIdx.x == 0 && sc[1] < n) *err_flag = 1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t nq = (nblk + 3) >> 2;
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    uint32_t w[4];
    load_owner4(owner, 4 * q, nblk, w);
    for (; q < nq; q += stride) {
        const int64_t p0 = 4 * q;
        uint32_t wn[4] = {0u, 0u, 0u, 0u};
        if (q + stride < nq) load_owner4(owner, 4 * (q + stride), nblk, wn);
        double v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e)
            v[e] = (w[e] != 0u && (uint64_t)(w[e] - 1) < n) ? xx[w[e] - 1] : 0.0;
        const int64_t ab64 = tri_row(p0);
        int ab = (int)ab64, cd = (int)(p0 - ab64 * (ab64 + 1) / 2);
        int i = (int)tri_row(ab), k = (int)tri_row(cd);
        int j = ab - i * (i + 1) / 2, l = cd - k * (k + 1) / 2;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int d = (i == j) + (k == l) + (ab == cd);
            double x = v[e];
            if (d >= 1) x = x + x;
            if (d >= 2) x = x + x;
            if (d >= 3) x = x + x;
            v[e] = x;
            // step to the next position: cd, then (k, l); past the diagonal, the next ab
            if (++l > k) { ++k; l = 0; }
            if (++cd > ab) {
                ++ab; cd = 0; k = 0; l = 0;
                if (++j > i) { ++i; j = 0; }
            }
        }
        if (VEC && p0 + 3 < nblk) {
            *(double2*)(out + p0) = make_double2(v[0], v[1]);
            *(double2*)(out + p0 + 2) = make_double2(v[2], v[3]);
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (p0 + e < nblk) out[p0 + e] = v[e];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) w[e] = wn[e];
    }
}

hipError_t launch_unpack_labels(const double* xx, const int32_t* labels, int64_t n_entries,
                                int n_ao, double* ao_packed, uint32_t* owner,
                                unsigned long long* sc, int* err_flag, hipStream_t st) {
    const int64_t pa = npairs(n_ao);
    const int64_t nblk = npairs(pa);
    hipLaunchKernelGGL(k_init_label_scalars, dim3(1), dim3(64), 0, st, sc, n_entries, err_flag);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nblk, 256)), dim3(256), 0, st, owner, nblk, 0u);
    if (n_entries > 0) {
        hipLaunchKernelGGL(k_claim_scan, dim3(grid_for(n_entries, kLabelThreads)),
                           dim3(kLabelThreads), 0, st, labels, n_entries, n_ao, owner, sc);
        hipLaunchKernelGGL(k_fix_clear, dim3(grid_for(nblk, 256)), dim3(256), 0, st, owner, nblk, sc);
        hipLaunchKernelGGL(k_fix_claim, dim3(grid_for(n_entries, 256)), dim3(256), 0, st, labels,
                           n_ao, owner, sc);
    }
    const dim3 rg(grid_for((nblk + 3) / 4, kLabelThreads));
    if ((uintptr_t)ao_packed % 16 == 0)
        hipLaunchKernelGGL(k_resolve<true>, rg, dim3(kLabelThreads), 0, st, xx, nblk, owner, sc,
                           err_flag, ao_packed);
    else
        hipLaunchKernelGGL(k_resolve<false>, rg, dim3(kLabelThreads), 0, st, xx, nblk, owner, sc,
                           err_flag, ao_packed);
    if (n_entries > 0)
        hipLaunchKernelGGL(k_noncanon_fix, dim3(grid_for(n_entries, 256, 2048)), dim3(256), 0, st,
                           xx, labels, n_ao, owner, sc, ao_packed);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------- input scan
// Finiteness and max |x| of a device copy of the caller's values, for the host entry
// points (which reject non-finite input before computing, and use max |x| in the
// split-FP16 error estimate).  s[0] = bit pattern of max |x| over the finite values (the
// order of non-negative doubles is the order of their bit patterns), s[1] = 1 if any
// value is NaN or infinite, s[2] = index of the first zero label (labels given) or n.
__global__ void k_scan_init(unsigned long long* s, int64_t n) {
    if (threadIdx.x == 0) { s[0] = 0ull; s[1] = 0ull; s[2] = (unsigned long long)n; }
}

__global__ __launch_bounds__(256) void k_first_zero(const int32_t* __restrict__ labels, int64_t n,
                                                    unsigned long long* s) {
    unsigned long long first = (unsigned long long)n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        if (labels[i] == 0) { first = (unsigned long long)i; break; }   // ascending per thread
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(first, o);
        first = x < first ? x : first;
    }
    if ((threadIdx.x & 63) == 0 && first < (unsigned long long)n) atomicMin(&s[2], first);
}

__global__ __launch_bounds__(256) void k_scan_finite(const double* __restrict__ p, int64_t n,
                                                     unsigned long long* s) {
    const int64_t end = (int64_t)s[2] < n ? (int64_t)s[2] : n;      // k_scan_init sets n
    unsigned long long m = 0ull, bad = 0ull;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t n2 = end >> 1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) {
        const double2 v = ((const double2*)p)[i];                    // p is 16-byte aligned
        const double a = fabs(v.x), b = fabs(v.y);
        if (!isfinite(a) || !isfinite(b)) bad = 1ull;
        else {
            const unsigned long long ba = (unsigned long long)__double_as_longlong(a);
            const unsigned long long bb = (unsigned long long)__double_as_longlong(b);
            m = ba > m ? ba : m;
            m = bb > m ? bb : m;
        }
    }
    if ((end & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const double a = fabs(p[end - 1]);
        if (!isfinite(a)) bad = 1ull;
        else {
            const unsigned long long ba = (unsigned long long)__double_as_longlong(a);
            m = ba > m ? ba : m;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(m, o);
        m = x > m ? x : m;
        bad |= __shfl_xor(bad, o);
    }
    if ((threadIdx.x & 63) == 0) {
        if (m) atomicMax(&s[0], m);
        if (bad) atomicOr(&s[1], 1ull);
    }
}

hipError_t launch_scan_input(const double* p, int64_t n, const int32_t* labels,
                             unsigned long long* s, hipStream_t st) {
    if (n < 0 || ((uintptr_t)p & 15)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scan_init, dim3(1), dim3(64), 0, st, s, n);
    if (n == 0) return hipGetLastError();
    if (labels)
        hipLaunchKernelGGL(k_first_zero, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, labels, n, s);
    hipLaunchKernelGGL(k_scan_finite, dim3(grid_for((n + 1) / 2, 256, 8192)), dim3(256), 0, st, p, n,
                       s);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------- expand
// The packed AO block stores each unordered pair-of-pairs once; the sandwich kernel
// wants row pq of the symmetric pair matrix contiguous.  64 x 64 tiles over the
// lower triangle: read the packed tile (rows contiguous), write it straight and
// transposed through LDS, both coalesced.  Row a of the pair matrix starts at vsq + a * ld
// (ld >= pa; the grouped expand pads rows to whole 128-byte lines).
constexpr int XT = kExpandTile;

__device__ __forceinline__ void expand_tile(const double* __restrict__ v, int64_t pa,
                                            double* __restrict__ vsq, int64_t ld, int64_t t) {
    __shared__ double tile[XT][XT + 1];
    const int64_t I = pair_row(t);
    const int64_t J = t - pair0(I, 0);
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;   // 64 x 4
    const int64_t a0 = I * XT, b0 = J * XT;
    for (int yy = ty; yy < XT; yy += 4) {
        int64_t a = a0 + yy, b = b0 + tx;
        double val = 0.0;
        if (a < pa && b < pa) val = v[pairsym(a, b)];
        tile[yy][tx] = val;
        if (a < pa && b < pa) vsq[a * ld + b] = val;
    }
    if (I == J) return;
    __syncthreads();
    for (int yy = ty; yy < XT; yy += 4) {
        int64_t b = b0 + yy, a = a0 + tx;
        if (a < pa && b < pa) vsq[b * ld + a] = tile[tx][yy];
    }
}

__global__ __launch_bounds__(256) void k_expand_packed(const double* __restrict__ v, int64_t pa,
                                                       double* __restrict__ vsq, int64_t t0) {
    expand_tile(v, pa, vsq, pa, t0 + blockIdx.x);
}

int64_t expand_tiles(int64_t pa) { return npairs((pa + XT - 1) / XT); }

hipError_t launch_expand_packed(const double* ao_packed, int64_t pa, double* vsq, hipStream_t st) {
    return launch_expand_rows(ao_packed, pa, vsq, 0, (pa + XT - 1) / XT, st);
}

// Tile rows [i0, i1): the tiles of rows 64 i0 .. 64 i1 - 1 of the pair matrix, which read
// only those rows of the packed block (a diagonal tile reads its own rows transposed).
hipError_t launch_expand_rows(const double* ao_packed, int64_t pa, double* vsq, int64_t i0,
                              int64_t i1, hipStream_t st) {
    const int64_t t0 = npairs(i0), nt = npairs(i1) - t0;
    if (nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_expand_packed, dim3((unsigned)nt), dim3(256), 0, st, ao_packed, pa, vsq,
                       t0);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------- transpose
using XTile = double[XT][XT + 1];

__device__ __forceinline__ void transpose_tile(XTile& tile, const double* __restrict__ src,
                                               int64_t rows, int64_t cols,
                                               double* __restrict__ dst, int64_t tiles_c,
                                               int64_t t) {
    const int64_t ti = t / tiles_c, tj = t - ti * tiles_c;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int64_t r0 = ti * XT, c0 = tj * XT;
    for (int yy = ty; yy < XT; yy += 4) {
        int64_t r = r0 + yy, c = c0 + tx;
        tile[yy][tx] = (r < rows && c < cols) ? src[r * cols + c] : 0.0;
    }
    __syncthreads();
    for (int yy = ty; yy < XT; yy += 4) {
        int64_t c = c0 + yy, r = r0 + tx;
        if (r < rows && c < cols) dst[c * rows + r] = tile[tx][yy];
    }
}

// The same tile with 16-byte accesses: a thread moves two adjacent doubles of a row on
// the way in and two adjacent doubles of a column on the way out.  Needs even rows and
// cols (every pair then starts on a 16-byte boundary of 16-byte-aligned buffers).
__device__ __forceinline__ void transpose_tile2(XTile& tile, const double* __restrict__ src,
                                                int64_t rows, int64_t cols,
                                                double* __restrict__ dst, int64_t tiles_c,
                                                int64_t t) {
    const int64_t ti = t / tiles_c, tj = t - ti * tiles_c;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;      // 32 x 8
    const int64_t r0 = ti * XT, c0 = tj * XT;
    for (int yy = ty; yy < XT; yy += 8) {
        const int64_t r = r0 + yy, c = c0 + 2 * tx;
        double2 v = make_double2(0.0, 0.0);
        if (r < rows && c < cols) v = *(const double2*)(src + r * cols + c);  // c + 1 < cols too
        tile[yy][2 * tx] = v.x;
        tile[yy][2 * tx + 1] = v.y;
    }
    __syncthreads();
    for (int yy = ty; yy < XT; yy += 8) {
        const int64_t c = c0 + yy, r = r0 + 2 * tx;
        if (r < rows && c < cols)
            *(double2*)(dst + c * rows + r) = make_double2(tile[2 * tx][yy], tile[2 * tx + 1][yy]);
    }
}

__global__ __launch_bounds__(256) void k_transpose(const double* __restrict__ src, int64_t rows,
                                                   int64_t cols, double* __restrict__ dst,
                                                   int64_t tiles_c) {
    __shared__ XTile tile;
    transpose_tile(tile, src, rows, cols, dst, tiles_c, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_transpose2(const double* __restrict__ src, int64_t rows,
                                                    int64_t cols, double* __restrict__ dst,
                                                    int64_t tiles_c) {
    __shared__ XTile tile;
    transpose_tile2(tile, src, rows, cols, dst, tiles_c, blockIdx.x);
}

int64_t transpose_tiles(int64_t rows, int64_t cols) {
    return ((rows + XT - 1) / XT) * ((cols + XT - 1) / XT);
}

hipError_t launch_transpose(const double* src, int64_t rows, int64_t cols, double* dst,
                            hipStream_t st) {
    const int64_t tc = (cols + XT - 1) / XT;
    const bool vec = (rows % 2 == 0) && (cols % 2 == 0) && ((uintptr_t)src % 16 == 0) &&
                     ((uintptr_t)dst % 16 == 0);
    if (vec)
        hipLaunchKernelGGL(k_transpose2, dim3((unsigned)transpose_tiles(rows, cols)), dim3(256), 0,
                           st, src, rows, cols, dst, tc);
    else
        hipLaunchKernelGGL(k_transpose, dim3((unsigned)transpose_tiles(rows, cols)), dim3(256), 0,
                           st, src, rows, cols, dst, tc);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------- C padding
// np x mp padded copy of C in 16-column blocks (cr_index; rows >= n, columns >= m zero),
// and the same values in the second-product order (cq_index) kCRows * mp doubles further on
__global__ void k_pad_c(const double* __restrict__ c, int ldc, int n, int m, int np, int mp,
                        double* __restrict__ cr) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)np * mp) return;
    int r = (int)(i / mp), a = (int)(i - (int64_t)r * mp);
    const double v = (r < n && a < m) ? c[r + (int64_t)a * ldc] : 0.0;
    cr[cr_index(r, a)] = v;
    cr[(int64_t)kCRows * mp + cq_index(r, a)] = v;
}

hipError_t launch_pad_c(const double* c, int ldc, int n, int m, int np, int mp, double* cr,
                        hipStream_t st) {
    int64_t tot = (int64_t)np * mp;
    hipLaunchKernelGGL(k_pad_c, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, c, ldc, n,
                       m, np, mp, cr);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------- grouped
// One launch over every transform of a scan: block / element indices are numbered
// transform after transform; the owning transform is found by binary search over the
// start offsets held in the (small, device-resident) table.
template <typename F>
__device__ __forceinline__ int find_geom(const LayoutGeom* __restrict__ g, int ngeom, int64_t x,
                                        F start) {
    int lo = 0, hi = ngeom;                       // start(g[0]) = 0 <= x
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start(g[mid]) <= x) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_group_expand(const LayoutGeom* __restrict__ g, int ngeom) {
    const int64_t t = blockIdx.x;
    const int k = find_geom(g, ngeom, t, [](const LayoutGeom& x) { return x.expand_start; });
    expand_tile(g[k].ao, g[k].pa, g[k].vsq, g[k].ldv, t - g[k].expand_start);
}

__global__ __launch_bounds__(256) void k_group_transpose(const LayoutGeom* __restrict__ g, int ngeom) {
    const int64_t t = blockIdx.x;
    const int k = find_geom(g, ngeom, t, [](const LayoutGeom& x) { return x.trans_start; });
    const LayoutGeom& G = g[k];
    const int64_t tc = (G.pm + XT - 1) / XT;
    // the per-geometry buffers come from hipMalloc (aligned); even extents allow the
    // 16-byte tile, and the choice is uniform over the work-group.  Both paths share one
    // LDS tile so the kernel keeps the occupancy of a single-variant transpose.
    __shared__ XTile tile;
    if (G.pa % 2 == 0 && G.pm % 2 == 0)
        transpose_tile2(tile, G.h, G.pa, G.pm, G.vsq, tc, t - G.trans_start);
    else
        transpose_tile(tile, G.h, G.pa, G.pm, G.vsq, tc, t - G.trans_start);
}

__global__ __launch_bounds__(256) void k_group_pad_c(const LayoutGeom* __restrict__ g, int ngeom,
                                                     int64_t total) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int k = find_geom(g, ngeom, i, [](const LayoutGeom& x) { return x.padc_start; });
    const LayoutGeom& G = g[k];
    const int64_t e = i - G.padc_start;
    const int r = (int)(e / G.mp), a = (int)(e - (int64_t)r * G.mp);
    const double v = (r < G.n && a < G.m) ? G.c[r + (int64_t)a * G.ldc] : 0.0;
    G.cr[cr_index(r, a)] = v;
    G.cr[(int64_t)kCRows * G.mp + cq_index(r, a)] = v;
}

hipError_t launch_group_expand(const LayoutGeom* g, int ngeom, int64_t total_tiles, hipStream_t st) {
    if (total_tiles <= 0 || total_tiles > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_group_expand, dim3((unsigned)total_tiles), dim3(256), 0, st, g, ngeom);
    return hipGetLastError();
}

hipError_t launch_group_transpose(const LayoutGeom* g, int ngeom, int64_t total_tiles,
                                  hipStream_t st) {
    if (total_tiles <= 0 || total_tiles > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_group_transpose, dim3((unsigned)total_tiles), dim3(256), 0, st, g, ngeom);
    return hipGetLastError();
}

hipError_t launch_group_pad_c(const LayoutGeom* g, int ngeom, int64_t total, hipStream_t st) {
    if (total <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_group_pad_c, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, g,
                       ngeom, total);
    return hipGetLastError();
}

}  // namespace cipsi
