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

__global__ __launch_bounds__(256) void k_expand_packed(const double* __restrict__ v, int64_t pa,
                                                       double* __restrict__ vsq) {
    __shared__ double tile[XT][XT + 1];
    const int64_t t = blockIdx.x;
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

hipError_t launch_expand_packed(const double* ao_packed, int64_t pa, double* vsq, hipStream_t st) {
    const int64_t nt = (pa + XT - 1) / XT;
    const int64_t ntiles = npairs(nt);
    hipLaunchKernelGGL(k_expand_packed, dim3((unsigned)ntiles), dim3(256), 0, st, ao_packed, pa, vsq);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------- transpose
__global__ __launch_bounds__(256) void k_transpose(const double* __restrict__ src, int64_t rows,
                                                   int64_t cols, double* __restrict__ dst,
                                                   int64_t tiles_c) {
    __shared__ double tile[XT][XT + 1];
    const int64_t t = blockIdx.x;
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

hipError_t launch_transpose(const double* src, int64_t rows, int64_t cols, double* dst,
                            hipStream_t st) {
    const int64_t tr = (rows + XT - 1) / XT, tc = (cols + XT - 1) / XT;
    hipLaunchKernelGGL(k_transpose, dim3((unsigned)(tr * tc)), dim3(256), 0, st, src, rows, cols,
                       dst, tc);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------- C padding
__global__ void k_pad_c(const double* __restrict__ c, int ldc, int n, int m, int np, int mp,
                        double* __restrict__ cr) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)np * mp) return;
    int r = (int)(i / mp), a = (int)(i - (int64_t)r * mp);
    cr[i] = (r < n && a < m) ? c[r + (int64_t)a * ldc] : 0.0;
}

hipError_t launch_pad_c(const double* c, int ldc, int n, int m, int np, int mp, double* cr,
                        hipStream_t st) {
    int64_t tot = (int64_t)np * mp;
    hipLaunchKernelGGL(k_pad_c, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, c, ldc, n,
                       m, np, mp, cr);
    return hipGetLastError();
}

}  // namespace cipsi
