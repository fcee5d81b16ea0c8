// HBM-bound layout kernels of the ijkl transform (gfx950).
//
//  unpack_labels  : pshf label stream -> packed AO block            (transd.f, C1 route)
//  expand_packed  : packed lower-triangular AO block -> full symmetric pair matrix
//  transpose      : H[pq][kl] -> H^T[kl][pq] between the two halves
//  pad_c          : C (column-major, ldc) -> zero-padded row-major copy used by the MFMA tiles
#include "common.h"

namespace cipsi {

// ----------------------------------------------------------------------------- fill
__global__ void k_fill_f64(double* __restrict__ p, int64_t n, double v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

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

hipError_t launch_zero(double* p, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n, 256)), dim3(256), 0, st, p, n, 0.0);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------- labels
// transd.f:85-97 (decode, first zero label ends the stream), 210-218 (doubling of
// the i==j, k==l and ij==kl entries that pshf/qout.f:73-77 halved), 369-372 (store
// at ij1(ij1-1)/2 + kl1).  The reference processes labels sequentially, so a later
// duplicate overwrites an earlier one: pass 1 elects the last writer per position
// with atomicMax on (index+1), pass 2 lets only that writer store.

__global__ void k_init_first_zero(int64_t* first_zero, int64_t n, int* err_flag) {
    if (threadIdx.x == 0 && blockIdx.x == 0) { *first_zero = n; *err_flag = 0; }
}

__global__ void k_find_first_zero(const int32_t* __restrict__ labels, int64_t n,
                                  int64_t* first_zero) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride)
        if (labels[i] == 0) atomicMin((unsigned long long*)first_zero, (unsigned long long)i);
}

__device__ inline bool decode_label(int32_t lab, int n_ao, int64_t nblk, int64_t* pos,
                                    int* dbl) {
    uint32_t u = (uint32_t)lab;
    int i = (int)((u >> 24) & 255u), j = (int)((u >> 16) & 255u);
    int k = (int)((u >> 8) & 255u), l = (int)(u & 255u);
    if (i < 1 || j < 1 || k < 1 || l < 1 || i > n_ao || j > n_ao || k > n_ao || l > n_ao)
        return false;
    int64_t ij1 = (int64_t)i * (i - 1) / 2 + j;
    int64_t kl1 = (int64_t)k * (k - 1) / 2 + l;
    int64_t p = ij1 * (ij1 - 1) / 2 + kl1 - 1;          // 0-based
    if (p < 0 || p >= nblk) return false;
    *pos = p;
    *dbl = (i == j) + (k == l) + (i == k && j == l);
    return true;
}

__global__ void k_claim_labels(const int32_t* __restrict__ labels, const int64_t* first_zero,
                               int n_ao, int64_t nblk, uint32_t* __restrict__ owner,
                               int* err_flag) {
    const int64_t n = *first_zero;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        int64_t pos;
        int dbl;
        if (!decode_label(labels[i], n_ao, nblk, &pos, &dbl)) { atomicOr(err_flag, 1); continue; }
        atomicMax(owner + pos, (uint32_t)(i + 1));
    }
}

__global__ void k_store_labels(const double* __restrict__ xx, const int32_t* __restrict__ labels,
                               const int64_t* first_zero, int n_ao, int64_t nblk,
                               const uint32_t* __restrict__ owner, double* __restrict__ out) {
    const int64_t n = *first_zero;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        int64_t pos;
        int dbl;
        if (!decode_label(labels[i], n_ao, nblk, &pos, &dbl)) continue;
        if (owner[pos] != (uint32_t)(i + 1)) continue;
        double v = xx[i];
        if (dbl >= 1) v = v + v;
        if (dbl >= 2) v = v + v;
        if (dbl >= 3) v = v + v;
        out[pos] = v;
    }
}

hipError_t launch_unpack_labels(const double* xx, const int32_t* labels, int64_t n_entries,
                                int n_ao, double* ao_packed, uint32_t* owner, int64_t* first_zero,
                                int* err_flag, hipStream_t st) {
    const int64_t pa = npairs(n_ao);
    const int64_t nblk = npairs(pa);
    hipLaunchKernelGGL(k_init_first_zero, dim3(1), dim3(64), 0, st, first_zero, n_entries, err_flag);
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(nblk, 256)), dim3(256), 0, st, ao_packed, nblk, 0.0);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nblk, 256)), dim3(256), 0, st, owner, nblk, 0u);
    if (n_entries > 0) {
        hipLaunchKernelGGL(k_find_first_zero, dim3(grid_for(n_entries, 256)), dim3(256), 0, st,
                           labels, n_entries, first_zero);
        hipLaunchKernelGGL(k_claim_labels, dim3(grid_for(n_entries, 256)), dim3(256), 0, st,
                           labels, first_zero, n_ao, nblk, owner, err_flag);
        hipLaunchKernelGGL(k_store_labels, dim3(grid_for(n_entries, 256)), dim3(256), 0, st,
                           xx, labels, first_zero, n_ao, nblk, owner, ao_packed);
    }
    return hipGetLastError();
}

// ----------------------------------------------------------------------------- expand
// The packed AO block stores each unordered pair-of-pairs once; the sandwich kernel
// wants row pq of the symmetric pair matrix contiguous.  64 x 64 tiles over the
// lower triangle: read the packed tile (rows contiguous), write it straight and
// transposed through LDS, both coalesced.
constexpr int XT = 64;

__device__ __forceinline__ void expand_tile(const double* __restrict__ v, int64_t pa,
                                            double* __restrict__ vsq, int64_t t) {
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
        if (a < pa && b < pa) vsq[a * pa + b] = val;
    }
    if (I == J) return;
    __syncthreads();
    for (int yy = ty; yy < XT; yy += 4) {
        int64_t b = b0 + yy, a = a0 + tx;
        if (a < pa && b < pa) vsq[b * pa + a] = tile[tx][yy];
    }
}

__global__ __launch_bounds__(256) void k_expand_packed(const double* __restrict__ v, int64_t pa,
                                                       double* __restrict__ vsq) {
    expand_tile(v, pa, vsq, blockIdx.x);
}

int64_t expand_tiles(int64_t pa) { return npairs((pa + XT - 1) / XT); }

hipError_t launch_expand_packed(const double* ao_packed, int64_t pa, double* vsq, hipStream_t st) {
    const int64_t ntiles = expand_tiles(pa);
    hipLaunchKernelGGL(k_expand_packed, dim3((unsigned)ntiles), dim3(256), 0, st, ao_packed, pa, vsq);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------- transpose
__device__ __forceinline__ void transpose_tile(const double* __restrict__ src, int64_t rows,
                                               int64_t cols, double* __restrict__ dst,
                                               int64_t tiles_c, int64_t t) {
    __shared__ double tile[XT][XT + 1];
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

__global__ __launch_bounds__(256) void k_transpose(const double* __restrict__ src, int64_t rows,
                                                   int64_t cols, double* __restrict__ dst,
                                                   int64_t tiles_c) {
    transpose_tile(src, rows, cols, dst, tiles_c, blockIdx.x);
}

int64_t transpose_tiles(int64_t rows, int64_t cols) {
    return ((rows + XT - 1) / XT) * ((cols + XT - 1) / XT);
}

hipError_t launch_transpose(const double* src, int64_t rows, int64_t cols, double* dst,
                            hipStream_t st) {
    const int64_t tc = (cols + XT - 1) / XT;
    hipLaunchKernelGGL(k_transpose, dim3((unsigned)transpose_tiles(rows, cols)), dim3(256), 0, st,
                       src, rows, cols, dst, tc);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------- C padding
// np x mp padded copy of C in 16-column blocks (cr_index; rows >= n, columns >= m zero)
__global__ void k_pad_c(const double* __restrict__ c, int ldc, int n, int m, int np, int mp,
                        double* __restrict__ cr) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)np * mp) return;
    int r = (int)(i / mp), a = (int)(i - (int64_t)r * mp);
    cr[cr_index(r, a)] = (r < n && a < m) ? c[r + (int64_t)a * ldc] : 0.0;
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
    expand_tile(g[k].ao, g[k].pa, g[k].vsq, t - g[k].expand_start);
}

__global__ __launch_bounds__(256) void k_group_transpose(const LayoutGeom* __restrict__ g, int ngeom) {
    const int64_t t = blockIdx.x;
    const int k = find_geom(g, ngeom, t, [](const LayoutGeom& x) { return x.trans_start; });
    transpose_tile(g[k].h, g[k].pa, g[k].pm, g[k].vsq, (g[k].pm + XT - 1) / XT,
                   t - g[k].trans_start);
}

__global__ __launch_bounds__(256) void k_group_pad_c(const LayoutGeom* __restrict__ g, int ngeom,
                                                     int64_t total) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int k = find_geom(g, ngeom, i, [](const LayoutGeom& x) { return x.padc_start; });
    const LayoutGeom& G = g[k];
    const int64_t e = i - G.padc_start;
    const int r = (int)(e / G.mp), a = (int)(e - (int64_t)r * G.mp);
    G.cr[cr_index(r, a)] = (r < G.n && a < G.m) ? G.c[r + (int64_t)a * G.ldc] : 0.0;
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
