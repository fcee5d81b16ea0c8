// FP64 MFMA half-transform ("sandwich") kernel for gfx950.
//
// One work-group = one symmetric matrix Z_b (an AO pair for half 1, an MO pair for
// half 2) times one tile of LT output columns:
//   Q1/Q3:  T = Z_b * C[:, c0:c0+LT]                (tpqkld.f:82-90 / tijkld.f:53-60)
//   Q2/Q4:  O(a, c) = sum_r C(r, a) T(r, c), a <= c  (tpqkld.f:91-97 / tijkld.f:61-76)
// Z_b panels (16 rows of the symmetric matrix) and the matching C rows are staged
// through LDS with register prefetch of the next panel; T stays in LDS for the
// second quarter.  Both quarters run on v_mfma_f64_16x16x4f64 (A: lane l holds
// A[l%16][l/16]; B: lane l holds B[l/16][l%16]; D: element q of lane l is
// D[l/16 + 4q][l%16]).
#include "common.h"

namespace cipsi {

constexpr int KC = 16;            // panel depth (rows of Z per LDS stage)

template <int LT, int NW>
struct HalfCfg {
    static constexpr int NT = NW * 64;
    static constexpr int NCB = LT / 16;
    static constexpr int RBMAX = (16 + NW - 1) / NW;      // np <= 256 -> <= 16 row blocks
    static constexpr int LDC = LT + 2;
    static constexpr int LDT = LT + 2;
    static constexpr int EPT_Z = (256 * KC + NT - 1) / NT;  // Z panel elements per thread
    static constexpr int EPT_C = (KC * LT + NT - 1) / NT;   // C panel elements per thread
};

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
            if (A.tile_start[mid] <= item) lo = mid; else hi = mid;
        }
        const int l = lo;
        const int nt = (l + lt) / lt;              // ceil((l+1)/lt)
        const int64_t rem = item - A.tile_start[l];
        b = pair0(l, 0) + rem / nt;
        tile = (int)(rem % nt);
        ncols = l + 1;
        limit = b + 1;
        dst_off = pair0(b, 0);
    }
}

template <int LT, int NW>
__global__ __launch_bounds__(NW * 64) void k_half_fp64(HalfArgs A) {
    using K = HalfCfg<LT, NW>;
    extern __shared__ double smem[];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int lm = lane & 15, lg = lane >> 4;
    const int np = A.np, n = A.n;
    const int LDZ = np + 2;

    // XCD-aware mapping: consecutive work items (same matrix, neighbouring tiles)
    // land on the same XCD so the Z_b row is fetched into one L2.
    const int64_t per_xcd = gridDim.x >> 3;
    const int64_t item = (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (item >= A.ntile_total) return;
    int64_t b, limit, dst_off;
    int tile, ncols;
    work_item(A, LT, item, b, tile, ncols, limit, dst_off);
    const int c0 = tile * LT;
    const int c1 = min(c0 + LT, ncols);
    const double* __restrict__ zsrc = A.src + b * A.src_ld;
    const double* __restrict__ cr = A.cr;
    const int ldm = A.ldm;

    double* zs_buf[2] = {smem, smem + (KC * LDZ + KC * K::LDC)};

    // ---- panel prefetch into registers: Z(r, s0+kk) for r < np, kk < KC; C(s0+kk, c0+c)
    double zst[K::EPT_Z];
    int zoff[K::EPT_Z];
    double cst[K::EPT_C];
    auto fetch = [&](int s0) {
        const int cntA = (np - s0) * KC;      // rows r >= s0: contiguous along s
#pragma unroll
        for (int t = 0; t < K::EPT_Z; ++t) {
            const int e = tid + t * K::NT;
            double v = 0.0;
            int off = -1;
            if (e < np * KC) {
                int r, kk;
                if (e < cntA) { r = s0 + e / KC; kk = e % KC; }
                else { const int e2 = e - cntA; kk = e2 / s0; r = e2 - kk * s0; }
                const int s = s0 + kk;
                if (r < n && s < n) v = zsrc[r >= s ? pair0(r, s) : pair0(s, r)];
                off = kk * LDZ + r;
            }
            zst[t] = v;
            zoff[t] = off;
        }
#pragma unroll
        for (int t = 0; t < K::EPT_C; ++t) {
            const int e = tid + t * K::NT;
            double v = 0.0;
            if (e < KC * LT) {
                const int kk = e / LT, c = e % LT;
                if (c0 + c < A.mp) v = cr[(int64_t)(s0 + kk) * ldm + c0 + c];
            }
            cst[t] = v;
        }
    };
    auto commit = [&](double* buf) {
        double* zs = buf;
        double* cs = buf + KC * LDZ;
#pragma unroll
        for (int t = 0; t < K::EPT_Z; ++t)
            if (zoff[t] >= 0) zs[zoff[t]] = zst[t];
#pragma unroll
        for (int t = 0; t < K::EPT_C; ++t) {
            const int e = tid + t * K::NT;
            if (e < KC * LT) cs[(e / LT) * K::LDC + (e % LT)] = cst[t];
        }
    };

    // ---- Q1: T = Z * C[:, c0:c0+LT]
    f64x4 acc[K::RBMAX][K::NCB];
#pragma unroll
    for (int i = 0; i < K::RBMAX; ++i)
#pragma unroll
        for (int j = 0; j < K::NCB; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

    const int nrb = np >> 4;
    const int nchunk = np / KC;
    fetch(0);
    commit(zs_buf[0]);
    __syncthreads();
    for (int ch = 0; ch < nchunk; ++ch) {
        const bool more = (ch + 1) < nchunk;
        if (more) fetch((ch + 1) * KC);
        const double* zs = zs_buf[ch & 1];
        const double* cs = zs + KC * LDZ;
#pragma unroll
        for (int k4 = 0; k4 < KC; k4 += 4) {
            double bf[K::NCB];
#pragma unroll
            for (int j = 0; j < K::NCB; ++j) bf[j] = cs[(k4 + lg) * K::LDC + j * 16 + lm];
#pragma unroll
            for (int i = 0; i < K::RBMAX; ++i) {
                const int rb = w + i * NW;
                if (rb < nrb) {
                    const double af = zs[(k4 + lg) * LDZ + rb * 16 + lm];
#pragma unroll
                    for (int j = 0; j < K::NCB; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, bf[j], acc[i][j], 0, 0, 0);
                }
            }
        }
        if (more) commit(zs_buf[(ch + 1) & 1]);
        __syncthreads();
    }

    // ---- T -> LDS (aliases the panel buffers; all reads finished at the barrier above)
    double* ts = smem;
#pragma unroll
    for (int i = 0; i < K::RBMAX; ++i) {
        const int rb = w + i * NW;
        if (rb < nrb) {
#pragma unroll
            for (int j = 0; j < K::NCB; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    ts[(rb * 16 + lg + 4 * q) * K::LDT + j * 16 + lm] = acc[i][j][q];
        }
    }
    __syncthreads();

    // ---- Q2: O(a, c) = sum_r C(r, a) T(r, c) for a <= c, c in [c0, c1)
    const int nab = (c1 - 1) / 16 + 1;         // row blocks that touch a <= c
    double* __restrict__ dst = A.dst + dst_off;
    for (int ab = w; ab < nab; ab += NW) {
        f64x4 o[K::NCB];
        bool act[K::NCB];
#pragma unroll
        for (int j = 0; j < K::NCB; ++j) {
            o[j] = f64x4{0.0, 0.0, 0.0, 0.0};
            const int cbeg = c0 + j * 16;
            act[j] = (cbeg < c1) && (cbeg + 15 >= ab * 16);
        }
        const double* ccol = cr + ab * 16 + lm;
        double an[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) an[u] = ccol[(int64_t)(u * 4 + lg) * ldm];
        for (int r0 = 0; r0 < np; r0 += 16) {
            double ac[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) ac[u] = an[u];
            if (r0 + 16 < np) {
#pragma unroll
                for (int u = 0; u < 4; ++u) an[u] = ccol[(int64_t)(r0 + 16 + u * 4 + lg) * ldm];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
#pragma unroll
                for (int j = 0; j < K::NCB; ++j) {
                    if (act[j]) {
                        const double bt = ts[(r0 + u * 4 + lg) * K::LDT + j * 16 + lm];
                        o[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[u], bt, o[j], 0, 0, 0);
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < K::NCB; ++j) {
            if (!act[j]) continue;
            const int c = c0 + j * 16 + lm;
            if (c >= c1) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int a = ab * 16 + lg + 4 * q;
                const int64_t p = pair0(c, a);
                if (a <= c && p < limit) dst[p] = o[j][q];
            }
        }
    }
}

template <int LT, int NW>
static hipError_t launch_lt(const HalfArgs& a, hipStream_t st) {
    using K = HalfCfg<LT, NW>;
    const int ldz = a.np + 2;
    const size_t panel = (size_t)(KC * ldz + KC * K::LDC);
    const size_t tsz = (size_t)a.np * K::LDT;
    const size_t shmem = sizeof(double) * (2 * panel > tsz ? 2 * panel : tsz);
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)k_half_fp64<LT, NW>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    int64_t g = (a.ntile_total + 7) / 8 * 8;
    if (g > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_half_fp64<LT, NW>), dim3((unsigned)g), dim3(K::NT), shmem, st, a);
    return hipGetLastError();
}

hipError_t launch_half_fp64(const HalfArgs& a, int lt, hipStream_t st) {
    if (a.np > 256 || (a.np & 15) || (a.mp & 15)) return hipErrorInvalidValue;
    switch (lt) {
        case 16: return launch_lt<16, 4>(a, st);
        case 32: return launch_lt<32, 4>(a, st);
        case 64: return launch_lt<64, 8>(a, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cipsi
