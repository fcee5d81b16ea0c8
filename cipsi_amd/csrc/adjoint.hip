// Energy E(C) = sum_pq h^MO_pq gamma_pq + 1/2 sum_pqrs Gamma_pqrs (pq|rs)_MO and its
// gradient dE/dC, with the half-transformed integrals recomputed (not stored).
//
// Conventions (0-based, pair0(a,b) = a(a+1)/2 + b, m(a) = 2 if a = pair(p,q) with p != q
// else 1):
//   gamma      full symmetric M x M (row-major)
//   Gamma      MO pair supermatrix, packed like the MO integrals: Gamma[pair0(a,b)], a>=b;
//              Gamma_pqrs = Gamma[pair(p,q), pair(r,s)] for every index order.
//   E2 = 1/2 sum_{a,b} m(a) m(b) Gamma_ab g_ab   (the full 4-index sum)
//   dE1/dC = 2 h C gamma
//   dE2/dC(mu,p) = 2 sum_{nu,q} C(nu,q) L[pair(mu,nu)][pair(p,q)],
//   L = H * Gh,  H[mn][b] = first half (tpqkld.f), Gh[b][a] = m(b) Gamma_ab.
// L is never formed.  With H_b(mu,nu) = H[pair(mu,nu)][b] and Gamma_b(q,p) =
// Gamma(pair(q,p), b), both symmetric, the sum regroups per MO pair b:
//   dE2/dC = 2 sum_b m(b) H_b C Gamma_b,
// two small products per b (2 N^2 M + 2 N M^2 flops) instead of the P_A x P_M x P_M
// product (2 P_A P_M^2).  All of it is FP64, on the matrix cores.
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/cipsi_ijkl.h"
#include "common.h"
#include "ctx.h"

using namespace cipsi;

namespace {

__device__ inline double pair_mult(int64_t a) {
    const int64_t p = pair_row(a);
    return (a - pair0(p, 0) == p) ? 1.0 : 2.0;
}

// Compensated (double-double) accumulation for the energy: error-free two-sum of the
// running sum with each term, the low parts summed separately (Knuth's TwoSum).
struct DD {
    double hi, lo;
};
__device__ inline void dd_add(DD& a, double x) {
#pragma clang fp contract(off)
    const double s = a.hi + x;
    const double bp = s - a.hi;
    const double err = (a.hi - (s - bp)) + (x - bp);
    a.hi = s;
    a.lo += err;
}
__device__ inline void dd_add(DD& a, const DD& b) {
#pragma clang fp contract(off)
    dd_add(a, b.hi);
    a.lo += b.lo;
}

// m(b) for every MO pair b, and the symmetrised one-particle density packed like h_MO:
// gs[pair0(p, q)] = (gamma(p,q) + gamma(q,p)) / 2 (= gamma(p,q) for a symmetric gamma,
// exactly), the matrix whose C-derivative term is 2 h C gs.
__global__ void k_pair_tables(const double* __restrict__ gamma, int m, int64_t pm,
                              double* __restrict__ mult, double* __restrict__ gs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pm) return;
    const int64_t p = pair_row(i), q = i - pair0(p, 0);
    mult[i] = p == q ? 1.0 : 2.0;
    gs[i] = 0.5 * (gamma[p * m + q] + gamma[q * m + p]);
}

// Per-block partial sums of w m(a) m(b) Gamma_ab g_ab over the packed index (w = 2 off
// the diagonal of the pair supermatrix), in double-double; a fixed-order final pass
// makes the result deterministic.  Block k takes the contiguous range
// [k L, (k + 1) L) of the packed index, L a multiple of RB; a thread steps through it RB
// elements at a time and carries its (row a, column b) along, so only its first
// element is decoded with a square root.
constexpr int RB = 256;
__device__ inline void block_reduce_dd(DD v, DD* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int off = RB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) dd_add(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
}

__global__ __launch_bounds__(RB) void k_energy2_partial(const double* __restrict__ g,
                                                        const double* __restrict__ gp,
                                                        const double* __restrict__ mult,
                                                        int64_t n, int64_t len,
                                                        double* __restrict__ part) {
    __shared__ DD red[RB];
    DD s{0.0, 0.0};
    const int64_t i0 = (int64_t)blockIdx.x * len + threadIdx.x;
    const int64_t iend = (int64_t)(blockIdx.x + 1) * len < n ? (int64_t)(blockIdx.x + 1) * len : n;
    if (i0 < iend) {
        int64_t a = pair_row(i0), b = i0 - pair0(a, 0);
        for (int64_t i = i0; i < iend; i += RB) {
            const double w = (a == b) ? 1.0 : 2.0;
            dd_add(s, w * mult[a] * mult[b] * gp[i] * g[i]);
            b += RB;
            while (b > a) {
                b -= a + 1;
                ++a;
            }
        }
    }
    block_reduce_dd(s, red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = red[0].hi;
        part[2 * blockIdx.x + 1] = red[0].lo;
    }
}

// energy = sum_pq hmo(pq) gamma(p,q) + 1/2 sum(part); one block, both sums reduced
// across its threads in a fixed order.
__global__ __launch_bounds__(RB) void k_energy_final(const double* __restrict__ part, int np_,
                                                     const double* __restrict__ hmo,
                                                     const double* __restrict__ gamma, int m,
                                                     double* __restrict__ energy) {
    __shared__ DD red[RB];
    DD s{0.0, 0.0};
    for (int i = threadIdx.x; i < m * m; i += RB) {
        const int p = i / m, q = i - p * m;
        dd_add(s, hmo[pairsym(p, q)] * gamma[i]);
    }
    block_reduce_dd(s, red);
    const DD e1 = red[0];
    __syncthreads();                                   // everyone has read red[0]
    DD s2{0.0, 0.0};
    for (int i = threadIdx.x; i < np_; i += RB) dd_add(s2, DD{part[2 * i], part[2 * i + 1]});
    block_reduce_dd(s2, red);
    if (threadIdx.x == 0) {
        DD e = e1;
        dd_add(e, DD{0.5 * red[0].hi, 0.5 * red[0].lo});
        *energy = e.hi + e.lo;
    }
}

// ---------------------------------------------------------------- dE2/dC on the matrix cores
// dE/dC = 2 sum_b m(b) H_b C Gamma_b + 2 h C gs.  The one-electron term is one more
// matrix b = P_M after the MO pairs, with H_b = h, Gamma_b = gs and m(b) = 1.  One wave
// per work item: a chunk of consecutive matrices b and one block of 16 AO rows mu (items are numbered chunk-major, so the waves of
// a work-group and of neighbouring work-groups share their matrices b in cache).  For
// each b of the chunk:
//   stage 1  U(q, mu) = sum_nu C(nu, q) H_b(nu, mu)               (U = (H_b C)^T)
//            A operand C from its zero-padded copy (16-column blocks), B operand H_b
//            from the packed row b of H^T; one accumulator per 16 q.
//   stage 2  G(mu, p) += m(b) sum_q U(q, mu) Gamma_b(q, p)
//            In the v_mfma_f64_16x16x4f64 D layout lane (lm, lg) holds element j of
//            block qb as U(16 qb + lg + 4j, 16 mb + lm): exactly A[m = mu][k = q] of the
//            K step (qb, j), so U stays in registers.  The B operand Gamma_b(q, p) is
//            read from row b of the expanded pair density (a packed symmetric M x M).
// BP matrices b are taken per pass and share stage 1's C loads.  m(b) (1 or 2, so the
// scaling is exact) multiplies the H_b operand as it is loaded: stage 2 then takes U
// from the accumulators as it is, with no vector arithmetic on it.  After the chunk the wave writes G (its 16 rows mu, every p)
// to part[chunk][mu][p]; k_grad_reduce adds the chunks in a fixed order, so the
// gradient is the same on every run.
// All operand loads are buffer loads over one matrix row (H_b, Gamma_b) or over the
// padded C: an index outside the row -- pair(nu, mu) with nu or mu >= N, pair(q, p) with
// q or p >= M -- is past the row's last element, and the hardware returns 0 for it.  A
// matrix b past the chunk's end gets an empty range and contributes nothing.
// Operands are prefetched two K groups ahead into register rings, with scheduling fences
// between the groups so that the compiler keeps the rings at their size.  The file is
// compiled with every MFMA operand in the ordinary vector registers
// (-amdgpu-mfma-vgpr-form): with accumulator registers the compiler copied U back to
// vector registers on every K step of stage 1.
typedef __amdgpu_buffer_rsrc_t Rsrc;

__device__ __forceinline__ Rsrc row_rsrc(const double* p, int64_t elems) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)(elems * 8), 0x00020000);
}
__device__ __forceinline__ double bload(Rsrc r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ unsigned packed_off(unsigned hi, unsigned lo) {
    return (((hi * (hi + 1)) >> 1) + lo) * 8;
}

template <int QB, int BP>
__global__ __launch_bounds__(256) void k_grad_mfma(const double* __restrict__ ht, int64_t pa,
                                                   const double* __restrict__ gam, int64_t pm,
                                                   const double* __restrict__ cr,
                                                   const double* __restrict__ h1,
                                                   const double* __restrict__ g1, int n, int m,
                                                   int nmb, int chunk, int64_t nitems,
                                                   double* __restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lm = lane & 15, lg = lane >> 4;
    const int64_t item = xcd_item(blockIdx.x) * 4 + w;
    if (item >= nitems) return;                        // whole wave; no barriers follow
    const int64_t ch = item / nmb;
    const int mb = (int)(item - ch * nmb);
    const int64_t b0 = ch * chunk, b1 = b0 + chunk < pm + 1 ? b0 + chunk : pm + 1;
    const unsigned mu = 16 * mb + lm;
    const int nk = 4 * nmb;                            // K steps over nu (rows padded to 16)
    // padded C: block qb holds kCRows rows of 16 doubles (32 KiB)
    const Rsrc crs = row_rsrc(cr, (int64_t)QB * kCRows * 16);
    const unsigned coff = (lg * 16 + lm) * 8;          // C(4k + lg, 16 qb + lm) - k, qb terms
    unsigned ap[QB];                                   // (p(p+1)/2 + lg) 8, p = 16 pb + lm
#pragma unroll
    for (int pb = 0; pb < QB; ++pb) {
        const unsigned p = 16 * pb + lm;
        ap[pb] = (((p * (p + 1)) >> 1) + lg) * 8;
    }
    f64x4 G[QB];
#pragma unroll
    for (int i = 0; i < QB; ++i) G[i] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int64_t b = b0; b < b1; b += BP) {
        Rsrc hr[BP], gr[BP];
        double mult[BP];
#pragma unroll
        for (int e = 0; e < BP; ++e) {
            const bool ok = b + e < b1;
            const int64_t bb = ok ? b + e : b;
            const bool one = bb == pm;                 // the one-electron term
            hr[e] = row_rsrc(one ? h1 : ht + bb * pa, ok ? pa : 0);
            gr[e] = row_rsrc(one ? g1 : gam + bb * pm, ok ? pm : 0);
            mult[e] = ok ? (one ? 1.0 : pair_mult(bb)) : 0.0;
        }
        f64x4 U[BP][QB];

        // stage 1: K step k covers nu = 4k + {0..3}; lane (lm, lg) supplies nu = 4k + lg.
        // A step past the last one loads only zeros or unused values (its offsets are past
        // the rows), so the two-step loop needs no guard.
        auto ld1 = [&](int k, double* av, double* hv) {
            const unsigned nu = 4 * k + lg;
            const unsigned ho = nu > mu ? packed_off(nu, mu) : packed_off(mu, nu);
#pragma unroll
            for (int e = 0; e < BP; ++e) hv[e] = bload(hr[e], ho, 0) * mult[e];
#pragma unroll
            for (int qb = 0; qb < QB; ++qb) av[qb] = bload(crs, coff + k * 512, qb * (kCRows * 128));
        };
        auto mm1 = [&](const double* av, const double* hv) {
#pragma unroll
            for (int qb = 0; qb < QB; ++qb)
#pragma unroll
                for (int e = 0; e < BP; ++e)
                    U[e][qb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[qb], hv[e], U[e][qb], 0, 0, 0);
        };
        // The first step starts U from a zero accumulator inside the MFMA, so U is an
        // accumulator value from its first definition and stays in the accumulator
        // registers through the loop (a zero initialised outside would be copied back
        // from them on every iteration).  nk = 4 nmb >= 4.
        double a0[QB], a1[QB], h0[BP], h1[BP];
        ld1(0, a0, h0);
        ld1(1, a1, h1);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int qb = 0; qb < QB; ++qb)
#pragma unroll
            for (int e = 0; e < BP; ++e)
                U[e][qb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[qb], h0[e],
                                                                f64x4{0.0, 0.0, 0.0, 0.0}, 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
        ld1(2, a0, h0);
        __builtin_amdgcn_sched_barrier(0);
        mm1(a1, h1);
        __builtin_amdgcn_sched_barrier(0);
        for (int k = 2; k < nk; k += 2) {
            ld1(k + 1, a1, h1);
            __builtin_amdgcn_sched_barrier(0);
            mm1(a0, h0);
            __builtin_amdgcn_sched_barrier(0);
            // the step after the last would read C rows np .. np + 3, past the padded
            // copy's last block when np = kCRows: skipped (k is the same in every lane)
            if (k + 2 < nk) ld1(k + 2, a0, h0);
            __builtin_amdgcn_sched_barrier(0);
            mm1(a1, h1);
            __builtin_amdgcn_sched_barrier(0);
        }

        // stage 2: group g = (e, qb, j) is the K step q = 16 qb + 4 j + {0..3} of matrix
        // e; lane (lm, lg) loads Gamma_b(16 qb + 4 j + lg, 16 pb + lm) for every pb.
        constexpr int NG = BP * QB * 4;
        constexpr int RD = 3;
        double ring[RD][QB];
        // Offsets of Gamma_b(q, p), q = 16 qb + 4 j + lg, p = 16 pb + lm, at pair(q, p):
        //   pb < qb:  (q(q+1)/2 + lm) 8 + 128 pb            bq = first term, per group
        //   pb > qb:  (p(p+1)/2 + lg) 8 + (16 qb + 4 j) 8   ap[pb] = first term, per lane
        //   pb = qb:  the first form where q >= p, else the second.
        // The constant parts go into the instruction's offset field, so the loads of a
        // group cost no vector arithmetic beyond bq and the diagonal select.
        auto ld2 = [&](int g, double* gv) {
            const int e = g / (QB * 4), qb = (g >> 2) % QB, j = g & 3;
            // bq and ap[pb] + cq depend only on the lane and the group; making the lane
            // index and ap opaque per group keeps the compiler from hoisting all of them
            // out of the loop over b (they would stay live in registers instead of going
            // into the offset field).
            unsigned lgv = lg;
            asm volatile("" : "+v"(lgv));
#pragma unroll
            for (int pb = 0; pb < QB; ++pb)
                if (pb >= qb) asm volatile("" : "+v"(ap[pb]));
            const unsigned q = 16 * qb + 4 * j + lgv;
            const unsigned bq = (((q * (q + 1)) >> 1) + lm) * 8;
            const unsigned cq = (16 * qb + 4 * j) * 8;
#pragma unroll
            for (int pb = 0; pb < QB; ++pb) {
                unsigned off;
                if (pb < qb) off = bq + 128 * pb;
                else if (pb > qb) off = ap[pb] + cq;
                else off = (4 * j + lgv >= (unsigned)lm) ? bq + 128 * pb : ap[pb] + cq;
                gv[pb] = bload(gr[e], off, 0);
            }
        };
        ld2(0, ring[0]);
        if (NG > 1) ld2(1, ring[1]);
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            if (g + 2 < NG) ld2(g + 2, ring[(g + 2) % RD]);
            __builtin_amdgcn_sched_barrier(0);
            const int e = g / (QB * 4), qb = (g >> 2) % QB, j = g & 3;
#pragma unroll
            for (int pb = 0; pb < QB; ++pb)
                G[pb] = __builtin_amdgcn_mfma_f64_16x16x4f64(U[e][qb][j], ring[g % RD][pb], G[pb], 0,
                                                             0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // lane (lm, lg), element j of block pb: G(16 mb + lg + 4j, 16 pb + lm)
    double* __restrict__ dst = part + ch * (int64_t)n * m;
#pragma unroll
    for (int pb = 0; pb < QB; ++pb) {
        const int p = 16 * pb + lm;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = 16 * mb + lg + 4 * j;
            if (r < n && p < m) dst[(int64_t)r * m + p] = G[pb][j];
        }
    }
}

// grad(mu, p) = 2 sum_chunks part[chunk][mu][p], written column-major (ld n).  64 outputs
// per work-group; wave s takes the chunks congruent to s mod 4, in four interleaved
// running sums (chunks s, s + 4, s + 8, s + 12 mod 16), and all partial sums are added in
// a fixed order.
__global__ __launch_bounds__(256) void k_grad_reduce(const double* __restrict__ part, int64_t nch,
                                                     int n, int m, double* __restrict__ grad) {
    __shared__ double red[4][64];
    const int t = threadIdx.x & 63, s = threadIdx.x >> 6;
    const int64_t nm = (int64_t)n * m;
    const int64_t o = (int64_t)blockIdx.x * 64 + t;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    if (o < nm) {
        int64_t c = s;
        for (; c + 12 < nch; c += 16) {
            a0 += part[c * nm + o];
            a1 += part[(c + 4) * nm + o];
            a2 += part[(c + 8) * nm + o];
            a3 += part[(c + 12) * nm + o];
        }
        for (; c < nch; c += 4) a0 += part[c * nm + o];
    }
    red[s][t] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    if (s != 0 || o >= nm) return;
    const int mu = (int)(o / m), p = (int)(o - (int64_t)mu * m);
    grad[(int64_t)p * n + mu] = 2.0 * ((red[0][t] + red[1][t]) + (red[2][t] + red[3][t]));
}

// Work split of k_grad_mfma: chunks of MO pairs sized so that a launch has about 4096
// waves (4 per SIMD), at least 2 and at most 64 matrices per chunk, a multiple of the
// pass width.  Two matrices per pass only while a wave still fits twice on a SIMD
// (up to 5 blocks of 16 MOs); beyond that the second U set would cost occupancy.
constexpr int kGradPairMaxQB = 5;
struct GradPlan {
    int nmb, qb, bp, chunk;
    int64_t nch, nitems, grid;
};

GradPlan grad_plan(int n, int m) {
    GradPlan g{};
    g.nmb = (n + 15) / 16;
    g.qb = (m + 15) / 16;
    g.bp = g.qb <= kGradPairMaxQB ? 2 : 1;
    const int64_t nb = npairs(m) + 1;                  // MO pairs and the one-electron term
    int64_t c = (nb * g.nmb + 4095) / 4096;
    c = c < 2 ? 2 : (c > 64 ? 64 : c);
    c = (c + g.bp - 1) / g.bp * g.bp;
    g.chunk = (int)c;
    g.nch = (nb + c - 1) / c;
    g.nitems = g.nch * g.nmb;
    g.grid = xcd_grid((g.nitems + 3) / 4);
    return g;
}

template <int QB>
hipError_t launch_grad_qb(const GradPlan& g, const double* ht, int64_t pa, const double* gam,
                          int64_t pm, const double* cr, const double* h1, const double* g1, int n,
                          int m, double* part, hipStream_t st) {
    if (g.bp == 2) {
        if constexpr (QB <= kGradPairMaxQB)
            hipLaunchKernelGGL((k_grad_mfma<QB, 2>), dim3((unsigned)g.grid), dim3(256), 0, st, ht,
                               pa, gam, pm, cr, h1, g1, n, m, g.nmb, g.chunk, g.nitems, part);
        else
            return hipErrorInvalidValue;
    } else {
        hipLaunchKernelGGL((k_grad_mfma<QB, 1>), dim3((unsigned)g.grid), dim3(256), 0, st, ht, pa,
                           gam, pm, cr, h1, g1, n, m, g.nmb, g.chunk, g.nitems, part);
    }
    return hipGetLastError();
}

hipError_t launch_grad(const GradPlan& g, const double* ht, int64_t pa, const double* gam,
                       int64_t pm, const double* cr, const double* h1, const double* g1, int n,
                       int m, double* part, hipStream_t st) {
    if (g.grid > 0x7fffffff || n > kCRows || npairs(n) * 8 > 0x7fffffff) return hipErrorInvalidValue;
    switch (g.qb) {
#define CIPSI_GRAD_CASE(Q) \
    case Q: return launch_grad_qb<Q>(g, ht, pa, gam, pm, cr, h1, g1, n, m, part, st);
        CIPSI_GRAD_CASE(1) CIPSI_GRAD_CASE(2) CIPSI_GRAD_CASE(3) CIPSI_GRAD_CASE(4)
        CIPSI_GRAD_CASE(5) CIPSI_GRAD_CASE(6) CIPSI_GRAD_CASE(7) CIPSI_GRAD_CASE(8)
        CIPSI_GRAD_CASE(9) CIPSI_GRAD_CASE(10) CIPSI_GRAD_CASE(11) CIPSI_GRAD_CASE(12)
        CIPSI_GRAD_CASE(13) CIPSI_GRAD_CASE(14) CIPSI_GRAD_CASE(15) CIPSI_GRAD_CASE(16)
#undef CIPSI_GRAD_CASE
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

static int energy_impl(cipsi_ijkl_ctx* ctx, int n, int m, const double* d_ao, const double* d_hp,
                       const double* d_c, int ldc, const double* d_gamma, const double* d_Gamma,
                       double* d_energy, double* d_grad) {
    const int64_t pa = npairs(n), pm = npairs(m), nout = npairs(pm);
    int rc;
    if ((rc = ctx_alloc(ctx, ctx->vsq, sizeof(double) * pa * pa))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->h, sizeof(double) * pa * pm))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->adj_g, sizeof(double) * nout))) return rc;
    const GradPlan gp = grad_plan(n, m);
    if ((rc = ctx_alloc(ctx, ctx->adj_small,
                        sizeof(double) * (3 * pm + (size_t)gp.nch * n * m + 4096))))
        return rc;
    double* vsq = (double*)ctx->vsq.ptr;
    double* h = (double*)ctx->h.ptr;
    double* g = (double*)ctx->adj_g.ptr;
    double* hmo = (double*)ctx->adj_small.ptr;
    double* mult = hmo + pm;                           // m(b) per MO pair
    double* gs = mult + pm;                            // symmetrised gamma, packed
    double* part = gs + pm;                            // energy partials (4096), then grad
    double* gpart = part + 4096;                       // partials (nch x n x m)
    double* cr;
    if ((rc = run_pad_c(ctx, n, m, d_c, ldc, &cr))) return rc;
    {
        StageTimer tm(ctx, "expand");
        if (launch_expand_packed(d_ao, pa, vsq, ctx->stream) != hipSuccess) return CIPSI_ERR_HIP;
    }
    // recompute the first half (tpqkld), stored as H^T (a row of P_A per MO pair) for the
    // second half and the gradient, then the second half for the energy
    if ((rc = run_half(ctx, 1, n, m, vsq, cr, h, CIPSI_PREC_FP64, 0, 1))) return rc;
    if ((rc = run_half(ctx, 2, n, m, h, cr, g, CIPSI_PREC_FP64, 0))) return rc;
    // h_MO = C^T h C (tmono) as a one-matrix sandwich
    {
        HalfArgs a{};
        a.src = d_hp;
        a.src_ld = pa;
        a.cr = cr;
        a.ldm = (m + 15) & ~15;
        a.dst = hmo;
        a.n = n;
        a.m = m;
        a.np = (n + 15) & ~15;
        a.mp = (m + 15) & ~15;
        a.half = 1;
        a.nb = 1;
        a.dst_ld = pm;
        a.ntile_total = (m + variant_columns(0) - 1) / variant_columns(0);
        if (launch_half_fp64(a, 0, ctx->stream) != hipSuccess) return CIPSI_ERR_HIP;
    }
    const int nblk = 1024;
    {
        StageTimer tm(ctx, "energy");
        hipLaunchKernelGGL(k_pair_tables, dim3((unsigned)((pm + 255) / 256)), dim3(256), 0,
                           ctx->stream, d_gamma, m, pm, mult, gs);
        const int64_t len = ((nout + nblk - 1) / nblk + RB - 1) / RB * RB;
        hipLaunchKernelGGL(k_energy2_partial, dim3(nblk), dim3(RB), 0, ctx->stream, g, d_Gamma,
                           mult, nout, len, part);
        hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(RB), 0, ctx->stream, part, nblk, hmo,
                           d_gamma, m, d_energy);
    }
    if (hipGetLastError() != hipSuccess) return CIPSI_ERR_HIP;
    if (!d_grad) return CIPSI_OK;

    // gradient: Gamma expanded to the full P_M x P_M (row b = Gamma_b packed), then
    // 2 sum_b m(b) H_b C Gamma_b + 2 h C gs from the rows of H^T, and the reduction
    if ((rc = ctx_alloc(ctx, ctx->adj_gh, sizeof(double) * pm * pm))) return rc;
    double* gfull = (double*)ctx->adj_gh.ptr;
    {
        StageTimer tm(ctx, "grad_gamma");
        if (launch_expand_packed(d_Gamma, pm, gfull, ctx->stream) != hipSuccess)
            return CIPSI_ERR_HIP;
    }
    {
        StageTimer tm(ctx, "grad_mfma");
        if (launch_grad(gp, h, pa, gfull, pm, cr, d_hp, gs, n, m, gpart, ctx->stream) !=
            hipSuccess)
            return CIPSI_ERR_HIP;
    }
    StageTimer tm(ctx, "grad_reduce");
    hipLaunchKernelGGL(k_grad_reduce, dim3((unsigned)((n * m + 63) / 64)), dim3(256), 0,
                       ctx->stream, gpart, gp.nch, n, m, d_grad);
    return hipGetLastError() == hipSuccess ? CIPSI_OK : CIPSI_ERR_HIP;
}

static int check_args(int n, int m, int ldc) {
    if (n < 1 || m < 1 || m > n || ldc < n) return CIPSI_ERR_ARG;
    if (n > 256) return CIPSI_ERR_UNSUPPORTED;
    return CIPSI_OK;
}

extern "C" {

int cipsi_ijkl_energy_grad_dev(cipsi_ijkl_ctx* ctx, int n_ao, int n_mo, const double* d_ao,
                               const double* d_h_packed, const double* d_c, int ldc,
                               const double* d_gamma, const double* d_Gamma_packed,
                               double* d_energy, double* d_grad) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_args(n_ao, n_mo, ldc))) return rc;
    if (!d_ao || !d_h_packed || !d_c || !d_gamma || !d_Gamma_packed || !d_energy)
        return CIPSI_ERR_ARG;
    return energy_impl(ctx, n_ao, n_mo, d_ao, d_h_packed, d_c, ldc, d_gamma, d_Gamma_packed,
                       d_energy, d_grad);
}

int cipsi_ijkl_energy_grad(cipsi_ijkl_ctx* ctx, int n_ao, int n_mo, const double* ao,
                           const double* h_packed, const double* c, int ldc, const double* gamma,
                           const double* Gamma_packed, double* energy, double* grad) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_args(n_ao, n_mo, ldc))) return rc;
    if (!ao || !h_packed || !c || !gamma || !Gamma_packed || !energy) return CIPSI_ERR_ARG;
    const int64_t pa = npairs(n_ao), pm = npairs(n_mo);
    const size_t s_ao = sizeof(double) * npairs(pa), s_h = sizeof(double) * pa,
                 s_c = sizeof(double) * (size_t)ldc * n_mo, s_g = sizeof(double) * n_mo * n_mo,
                 s_G = sizeof(double) * npairs(pm), s_grad = sizeof(double) * (size_t)n_ao * n_mo;
    if ((rc = ctx_alloc(ctx, ctx->io_in, s_ao))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->io_in2, s_h + s_c + s_g + s_G + 64))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->io_out, s_grad + 64))) return rc;
    char* b2 = (char*)ctx->io_in2.ptr;
    double* d_h = (double*)b2;
    double* d_c = (double*)(b2 + s_h);
    double* d_gam = (double*)(b2 + s_h + s_c);
    double* d_Gam = (double*)(b2 + s_h + s_c + s_g);
    double* d_grad = (double*)ctx->io_out.ptr;
    double* d_e = (double*)((char*)ctx->io_out.ptr + s_grad);
    hipStream_t st = ctx->stream;
    // the two large arrays (AO block, pair density) go through the pinned bounce slots
    if ((rc = ctx_copy_h2d(ctx, ctx->io_in.ptr, ao, s_ao))) return rc;
    if ((rc = ctx_copy_h2d(ctx, d_Gam, Gamma_packed, s_G))) return rc;
    if (hipMemcpyAsync(d_h, h_packed, s_h, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_c, c, s_c, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_gam, gamma, s_g, hipMemcpyHostToDevice, st) != hipSuccess)
        return CIPSI_ERR_HIP;
    if ((rc = energy_impl(ctx, n_ao, n_mo, (const double*)ctx->io_in.ptr, d_h, d_c, ldc, d_gam,
                          d_Gam, d_e, grad ? d_grad : nullptr)))
        return rc;
    if (hipMemcpyAsync(energy, d_e, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
        return CIPSI_ERR_HIP;
    if (grad && hipMemcpyAsync(grad, d_grad, s_grad, hipMemcpyDeviceToHost, st) != hipSuccess)
        return CIPSI_ERR_HIP;
    return hipStreamSynchronize(st) == hipSuccess ? CIPSI_OK : CIPSI_ERR_HIP;
}

}  // extern "C"
