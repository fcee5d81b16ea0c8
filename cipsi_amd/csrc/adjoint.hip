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
// The P_A x P_M x P_M product L = H Gh is a plain GEMM (rocBLAS dgemm); everything
// else is here, in FP64.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

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

// Gh[b][a] = m(b) Gamma(a, b), full P_M x P_M row-major.
__global__ void k_gamma_hat(const double* __restrict__ gp, int64_t pm, double* __restrict__ gh) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < pm * pm; i += stride) {
        const int64_t b = i / pm, a = i - b * pm;
        gh[i] = pair_mult(b) * gp[pairsym(a, b)];
    }
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

// Per-block partial sums of w m(a) m(b) Gamma_ab g_ab over the packed index (w = 2 off
// the diagonal of the pair supermatrix), in double-double; a fixed-order final pass
// makes the result deterministic.
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
                                                        int64_t n, double* __restrict__ part) {
    __shared__ DD red[RB];
    DD s{0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
        const int64_t a = pair_row(i), b = i - pair0(a, 0);
        const double w = (a == b) ? 1.0 : 2.0;
        dd_add(s, w * pair_mult(a) * pair_mult(b) * gp[i] * g[i]);
    }
    block_reduce_dd(s, red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = red[0].hi;
        part[2 * blockIdx.x + 1] = red[0].lo;
    }
}

// energy = sum_pq hmo(pq) gamma(p,q) + 1/2 sum(part); one block.
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
    if (threadIdx.x == 0) {
        DD e2{0.0, 0.0};
        for (int i = 0; i < np_; ++i) dd_add(e2, DD{part[2 * i], part[2 * i + 1]});
        DD e = red[0];
        dd_add(e, DD{0.5 * e2.hi, 0.5 * e2.lo});
        *energy = e.hi + e.lo;
    }
}

// hc(mu, q) = sum_nu h(mu, nu) C(nu, q)   (N x M, column-major, ld n)
__global__ void k_hc(const double* __restrict__ hp, const double* __restrict__ c, int ldc, int n,
                     int m, double* __restrict__ hc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * m) return;
    const int q = i / n, mu = i - q * n;
    double s = 0.0;
    for (int nu = 0; nu < n; ++nu) s += hp[pairsym(mu, nu)] * c[(int64_t)q * ldc + nu];
    hc[i] = s;
}

// grad(mu, p) = 2 sum_q hc(mu,q) gamma(q,p) + 2 sum_{nu,q} C(nu,q) L[pair(mu,nu)][pair(p,q)]
// Row r = pair(mu, nu) of L, read as the symmetric M x M matrix M_r(p, q) =
// L[r][pair(p,q)], enters grad(mu, .) as M_r C(nu, .)^T.  One work-group per (mu, chunk
// of kGradNu values of nu) streams those rows once each: thread p forms sum_q M_r(p,q)
// C(nu,q), reading its own contiguous piece (q <= p) and then, coalesced with the other
// threads, the column below the diagonal (q > p).  C(nu, .) is staged in LDS.  Partial
// sums per chunk go to part[chunk][mu][p]; k_grad_sum adds them in chunk order (so the
// result is the same on every run) together with the one-electron term.
constexpr int kGradNu = 8;

__global__ __launch_bounds__(256) void k_grad_rows(const double* __restrict__ L, int64_t pm,
                                                   const double* __restrict__ c, int ldc, int n,
                                                   int m, double* __restrict__ part) {
    __shared__ double cs[256];
    const int mu = blockIdx.x, ch = blockIdx.y;
    const int p = threadIdx.x;
    const int nu0 = ch * kGradNu, nu1 = min(n, nu0 + kGradNu);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};               // q mod 4: independent FMA chains
    for (int nu = nu0; nu < nu1; ++nu) {
        __syncthreads();                               // previous nu's reads of cs done
        if (p < m) cs[p] = c[(int64_t)p * ldc + nu];
        __syncthreads();
        if (p < m) {
            const double* __restrict__ row = L + pairsym(mu, nu) * pm;
            const double* __restrict__ rp = row + (int64_t)p * (p + 1) / 2;
            int q = 0;
            for (; q + 3 <= p; q += 4) {
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[e] += rp[q + e] * cs[q + e];
            }
            for (; q <= p; ++q) acc[q & 3] += rp[q] * cs[q];
            for (q = p + 1; q + 3 < m; q += 4) {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    acc[(q + e) & 3] += row[(int64_t)(q + e) * (q + e + 1) / 2 + p] * cs[q + e];
            }
            for (; q < m; ++q) acc[q & 3] += row[(int64_t)q * (q + 1) / 2 + p] * cs[q];
        }
    }
    if (p < m) part[((int64_t)ch * n + mu) * m + p] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

__global__ __launch_bounds__(256) void k_grad_sum(const double* __restrict__ part, int nch,
                                                  const double* __restrict__ hc,
                                                  const double* __restrict__ gamma, int n, int m,
                                                  double* __restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * m) return;
    const int mu = i / m, p = i - mu * m;
    double s = 0.0;
    for (int ch = 0; ch < nch; ++ch) s += part[((int64_t)ch * n + mu) * m + p];
    for (int q = 0; q < m; ++q) s += hc[(int64_t)q * n + mu] * gamma[(int64_t)q * m + p];
    grad[(int64_t)p * n + mu] = 2.0 * s;
}

}  // namespace

int ctx_rocblas(cipsi_ijkl_ctx* ctx);

static int energy_impl(cipsi_ijkl_ctx* ctx, int n, int m, const double* d_ao, const double* d_hp,
                       const double* d_c, int ldc, const double* d_gamma, const double* d_Gamma,
                       double* d_energy, double* d_grad) {
    const int64_t pa = npairs(n), pm = npairs(m), nout = npairs(pm);
    int rc;
    if ((rc = ctx_alloc(ctx, ctx->vsq, sizeof(double) * pa * pa))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->h, sizeof(double) * pa * pm))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->adj_g, sizeof(double) * nout))) return rc;
    const int nch = (n + kGradNu - 1) / kGradNu;
    if ((rc = ctx_alloc(ctx, ctx->adj_small,
                        sizeof(double) * (pm + (2 + (size_t)nch) * n * m + 4096))))
        return rc;
    double* vsq = (double*)ctx->vsq.ptr;
    double* h = (double*)ctx->h.ptr;
    double* g = (double*)ctx->adj_g.ptr;
    double* hmo = (double*)ctx->adj_small.ptr;
    double* hc = hmo + pm;
    double* part = hc + (size_t)n * m;                // energy partials (4096), then grad
    double* gpart = part + 4096;                       // partials (nch x n x m)
    double* cr;
    if ((rc = run_pad_c(ctx, n, m, d_c, ldc, &cr))) return rc;
    {
        StageTimer tm(ctx, "expand");
        if (launch_expand_packed(d_ao, pa, vsq, ctx->stream) != hipSuccess) return CIPSI_ERR_HIP;
    }
    // recompute the first half (tpqkld), then the second half for the energy
    if ((rc = run_half(ctx, 1, n, m, vsq, cr, h, CIPSI_PREC_FP64, 0))) return rc;
    {
        StageTimer tm(ctx, "transpose");
        if (launch_transpose(h, pa, pm, vsq, ctx->stream) != hipSuccess) return CIPSI_ERR_HIP;
    }
    if ((rc = run_half(ctx, 2, n, m, vsq, cr, g, CIPSI_PREC_FP64, 0))) return rc;
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
        hipLaunchKernelGGL(k_energy2_partial, dim3(nblk), dim3(RB), 0, ctx->stream, g, d_Gamma,
                           nout, part);
        hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(RB), 0, ctx->stream, part, nblk, hmo,
                           d_gamma, m, d_energy);
    }
    if (hipGetLastError() != hipSuccess) return CIPSI_ERR_HIP;
    if (!d_grad) return CIPSI_OK;

    // gradient: L = H Gh (rocBLAS, column-major view: L^T = Gh^T H^T), then the contraction
    if ((rc = ctx_alloc(ctx, ctx->adj_gh, sizeof(double) * pm * pm))) return rc;
    double* gh = (double*)ctx->adj_gh.ptr;
    double* L = vsq;                                   // P_A x P_M fits in the P_A^2 buffer
    if ((rc = ctx_rocblas(ctx))) return rc;
    {
        StageTimer tm(ctx, "grad_gemm");
        hipLaunchKernelGGL(k_gamma_hat, dim3(4096), dim3(256), 0, ctx->stream, d_Gamma, pm, gh);
        const double one = 1.0, zero = 0.0;
        if (rocblas_dgemm((rocblas_handle)ctx->blas, rocblas_operation_none,
                          rocblas_operation_none, (rocblas_int)pm, (rocblas_int)pa,
                          (rocblas_int)pm, &one, gh, (rocblas_int)pm, h, (rocblas_int)pm, &zero,
                          L, (rocblas_int)pm) != rocblas_status_success)
            return CIPSI_ERR_HIP;
    }
    StageTimer tm(ctx, "grad_contract");
    hipLaunchKernelGGL(k_hc, dim3((n * m + 255) / 256), dim3(256), 0, ctx->stream, d_hp, d_c, ldc,
                       n, m, hc);
    hipLaunchKernelGGL(k_grad_rows, dim3(n, nch), dim3(256), 0, ctx->stream, L, pm, d_c, ldc, n, m,
                       gpart);
    hipLaunchKernelGGL(k_grad_sum, dim3((n * m + 255) / 256), dim3(256), 0, ctx->stream, gpart,
                       nch, hc, d_gamma, n, m, d_grad);
    return hipGetLastError() == hipSuccess ? CIPSI_OK : CIPSI_ERR_HIP;
}

int ctx_rocblas(cipsi_ijkl_ctx* ctx) {
    if (!ctx->blas) {
        rocblas_handle hdl;
        if (rocblas_create_handle(&hdl) != rocblas_status_success) return CIPSI_ERR_HIP;
        ctx->blas = hdl;
    }
    if (rocblas_set_stream((rocblas_handle)ctx->blas, ctx->stream) != rocblas_status_success)
        return CIPSI_ERR_HIP;
    return CIPSI_OK;
}

void ctx_rocblas_destroy(cipsi_ijkl_ctx* ctx) {
    if (ctx->blas) (void)rocblas_destroy_handle((rocblas_handle)ctx->blas);
    ctx->blas = nullptr;
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
