// Integral consumers of the packed MO stream, on the device.
//
//   gather   ai(i,j,k,l) for a batch of integral labels: the random access that the
//            downstream programs make into the MO integrals (fock/ai.f:12-35, C1 route:
//            pair indices, larger pair first; cip/ai.f:1-42 for the same lookup).
//   fock     the closed-shell reference energy, the MO Fock operator and the J / K
//            diagonals of CIPSI_sources/fock/fock.f:80-127.
//   hdig     cip's determinant energies relative to the reference (cip/hdig.f:40-80),
//   hntd     and its off-diagonal Slater-Condon elements (cip/hntd.f:36-168), for
//            determinants held as excitations of the closed-shell reference.
//
// Conventions (0-based here, the reference is 1-based):
//   mo[pair0(max(a,b), min(a,b))], a = pair(i,j), b = pair(k,l)   the packed MO stream
//   hmo[pair0(i,j)], i >= j                                        h_MO (tmono.f:47-54)
//   label = i<<24 | j<<16 | k<<8 | l, 1-based 8-bit fields       (ijkl/iword_func.f:2-52)
//
// The Fock elements and the energy are summed in exactly the order of fock.f, one element
// per thread, with multiply-add contraction off: on the same MO integrals the results are
// the bits of a build of fock.f without FMA.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/cipsi_ijkl.h"
#include "common.h"
#include "ctx.h"

using namespace cipsi;

namespace {

// fock.f:3420: elements below this magnitude are set to zero.
constexpr double kFockThreshold = 1.0e-7;
constexpr int kEnergyThreads = 256;
constexpr int kMaxMo = 256;       // n_mo cap of the library (8-bit labels cap gather at 255)
constexpr int kEnergyStage = 8192;   // doubles of tt staged in LDS per pass (64 KiB)

__device__ inline double ai(const double* __restrict__ mo, int i, int j, int k, int l) {
    return mo[pairsym(pairsym(i, j), pairsym(k, l))];
}

// One thread per label; out-of-range fields give NaN and are counted.
__global__ void k_gather(const double* __restrict__ mo, int m, const int32_t* __restrict__ lab,
                         int64_t n, double* __restrict__ out,
                         unsigned long long* __restrict__ n_bad) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        const uint32_t u = (uint32_t)lab[q];
        const int i = (int)(u >> 24), j = (int)((u >> 16) & 255u), k = (int)((u >> 8) & 255u),
                  l = (int)(u & 255u);
        const bool ok = i >= 1 && j >= 1 && k >= 1 && l >= 1 && i <= m && j <= m && k <= m &&
                        l <= m;
        if (ok) {
            out[q] = ai(mo, i - 1, j - 1, k - 1, l - 1);
        } else {
            out[q] = __builtin_nan("");
            atomicAdd(n_bad, 1ull);
        }
    }
}

// fock.f:81-127 per element (i >= j), in the reference's order of additions:
//   i < noc, j < i   f = h; for k < noc: f = f + (2 (ij|kk) - (ik|jk))           (:87-96)
//   i < noc, j = i   f = h + (ii|ii); then tt(i,k) = 2(ii|kk) - (ik|ik) for k < i, then
//                    for k = i+1 .. noc-1 (added when the outer loop reaches k, :85-93)
//   i >= noc         tt = 0; for k < noc: tt = tt + (2 (ij|kk) - (ik|jk)); f = h + tt  (:102-109)
// then |f| < 1e-7 -> 0 (:3420), and J = (ii|jj), K = (ij|ij) (:110-114).
__global__ void k_fock_elements(const double* __restrict__ mo, const double* __restrict__ hmo,
                                int m, int noc, double* __restrict__ fock,
                                double* __restrict__ jint, double* __restrict__ kint) {
#pragma clang fp contract(off)
    const int64_t pm = pair0(m, 0);
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= pm) return;
    const int i = (int)pair_row(x);
    const int j = (int)(x - pair0(i, 0));
    const double h = hmo[x];
    double f;
    if (i >= noc) {
        double tt = 0.0;
#pragma unroll 4
        for (int k = 0; k < noc; ++k) tt = tt + (2.0 * ai(mo, i, j, k, k) - ai(mo, i, k, j, k));
        f = h + tt;
    } else if (j < i) {
        f = h;
#pragma unroll 4
        for (int k = 0; k < noc; ++k) f = f + (ai(mo, i, j, k, k) * 2.0 - ai(mo, i, k, j, k));
    } else {
        f = h + ai(mo, i, i, i, i);
        for (int k = 0; k < noc; ++k) {
            if (k == i) continue;
            double tt = ai(mo, i, i, k, k);
            const double tk = ai(mo, i, k, i, k);
            tt = tt + tt - tk;
            f = f + tt;
        }
    }
    fock[x] = fabs(f) < kFockThreshold ? 0.0 : f;
    jint[x] = ai(mo, i, i, j, j);
    kint[x] = ai(mo, i, j, i, j);
}

// fock.f:80-99: val, the closed-shell electronic energy (without the nuclear term), one
// work-group, after k_fock_elements.  The terms tt(i,k) = 2 (ii|kk) - (ik|ik) are formed
// from the J and K rows just written (the same stored integrals ai() returns), staged in
// LDS a block of rows at a time, and added by one lane in the reference's order:
// 2 h_ii, (ii|ii), then tt(i,k) twice for each k < i.
__global__ void __launch_bounds__(kEnergyThreads)
k_fock_energy(const double* __restrict__ hmo, const double* __restrict__ jint,
              const double* __restrict__ kint, int noc, double* __restrict__ energy) {
#pragma clang fp contract(off)
    __shared__ double tt[kEnergyStage];
    __shared__ double hd[kMaxMo];
    double val = 0.0;
    int i0 = 0;
    while (i0 < noc) {
        const int64_t base = pair0(i0, 0);
        int i1 = i0 + 1;
        while (i1 < noc && pair0(i1 + 1, 0) - base <= kEnergyStage) ++i1;
        const int cnt = (int)(pair0(i1, 0) - base);      // rows i0 .. i1-1, k <= i
        for (int x = threadIdx.x; x < cnt; x += blockDim.x) {
            const double jv = jint[base + x], kv = kint[base + x];
            tt[x] = jv + jv - kv;                          // overwritten below on the diagonal
        }
        __syncthreads();
        for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
            const int64_t d = pair0(i, i);
            tt[d - base] = jint[d];                        // (ii|ii)
            hd[i - i0] = hmo[d];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int i = i0; i < i1; ++i) {
                const int r = (int)(pair0(i, 0) - base);
                val = val + hd[i - i0] * 2.0;
                val = val + tt[r + i];
                int k = 0;
                for (; k + 8 <= i; k += 8) {
                    double t[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) t[u] = tt[r + k + u];
#pragma unroll
                    for (int u = 0; u < 8; ++u) val = val + t[u] + t[u];
                }
                for (; k < i; ++k) val = val + tt[r + k] + tt[r + k];
            }
        }
        __syncthreads();
        i0 = i1;
    }
    if (threadIdx.x == 0) *energy = val;
}

int cons_scalars(cipsi_ijkl_ctx* ctx, unsigned long long** n_bad, double** energy) {
    int rc = ctx_alloc(ctx, ctx->cons, 64);
    if (rc) return rc;
    *n_bad = (unsigned long long*)ctx->cons.ptr;
    *energy = (double*)((char*)ctx->cons.ptr + 16);
    return CIPSI_OK;
}

int gather_impl(cipsi_ijkl_ctx* ctx, int m, const double* d_mo, const int32_t* d_labels,
                int64_t n, double* d_out) {
    unsigned long long* n_bad;
    double* e;
    int rc = cons_scalars(ctx, &n_bad, &e);
    if (rc) return rc;
    if (hipMemsetAsync(n_bad, 0, sizeof(unsigned long long), ctx->stream) != hipSuccess)
        return CIPSI_ERR_HIP;
    if (n > 0) {
        StageTimer tm(ctx, "gather");
        const int64_t blocks = (n + 255) / 256;
        const unsigned grid = (unsigned)(blocks < 65536 ? blocks : 65536);
        hipLaunchKernelGGL(k_gather, dim3(grid), dim3(256), 0, ctx->stream, d_mo, m, d_labels, n,
                           d_out, n_bad);
        if (hipGetLastError() != hipSuccess) return CIPSI_ERR_HIP;
    }
    unsigned long long bad = 0;
    if (hipMemcpyAsync(&bad, n_bad, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return CIPSI_ERR_HIP;
    if (bad) {
        ctx->last_error = "gather: " + std::to_string(bad) + " labels with a field outside 1..n_mo";
        return CIPSI_ERR_LABEL;
    }
    return CIPSI_OK;
}

int fock_impl(cipsi_ijkl_ctx* ctx, int m, int noc, const double* d_mo, const double* d_hmo,
              double* d_energy, double* d_fock, double* d_j, double* d_k) {
    StageTimer tm(ctx, "fock");
    const int64_t pm = pair0(m, 0);
    hipLaunchKernelGGL(k_fock_elements, dim3((unsigned)((pm + 63) / 64)), dim3(64), 0, ctx->stream,
                       d_mo, d_hmo, m, noc, d_fock, d_j, d_k);
    hipLaunchKernelGGL(k_fock_energy, dim3(1), dim3(kEnergyThreads), 0, ctx->stream, d_hmo, d_j,
                       d_k, noc, d_energy);
    return hipGetLastError() == hipSuccess ? CIPSI_OK : CIPSI_ERR_HIP;
}

int check_fock_args(int m, int noc) {
    if (m < 1 || noc < 0 || noc > m) return CIPSI_ERR_ARG;
    if (m > kMaxMo) return CIPSI_ERR_UNSUPPORTED;
    return CIPSI_OK;
}


// ----------------------------------------------------------------------------- Slater-Condon
// cip's determinant arrays (common /spo/), all numbers 1-based as there: determinant d
// has ne[d] excitations, holes trou[nd[d] + t] and particles part[nd[d] + t] (spin-
// orbital numbers); spin-orbital s has orbital iorb[s - 1] and spin ispin[s - 1].
struct DetArrays {
    const int* ne;
    const int* nd;
    const int* trou;
    const int* part;
    const int* iorb;
    const int* ispin;
};
constexpr int kMaxExc = 10;           // cip's ita(20) / itb(20) hold 2 x 10 entries

// 0-based packed index of the 1-based pair (a, b) as cip's num(a) + b, not reordered.
__device__ inline int64_t tri1(int64_t a, int64_t b) { return a * (a - 1) / 2 + b - 1; }
__device__ inline int64_t tri1s(int64_t a, int64_t b) { return a >= b ? tri1(a, b) : tri1(b, a); }
__device__ inline double ai1(const double* __restrict__ mo, int a, int b, int c, int e) {
    return ai(mo, a - 1, b - 1, c - 1, e - 1);
}

// Every determinant and spin-orbital in range, at most kMaxExc excitations; counts the
// violations.
__global__ void k_sc_validate(DetArrays D, int n_det, int64_t n_exc, int n_so, int n_mo,
                              unsigned long long* __restrict__ n_bad) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t s = t0; s < n_so; s += stride) {
        const int o = D.iorb[s], sp = D.ispin[s];
        if (o < 1 || o > n_mo || sp < 1 || sp > 2) atomicAdd(n_bad, 1ull);
    }
    for (int64_t d = t0; d < n_det; d += stride) {
        const int e = D.ne[d], o = D.nd[d];
        bool ok = e >= 0 && e <= kMaxExc && o >= 0 && (int64_t)o + e <= n_exc;
        for (int t = 0; ok && t < e; ++t) {
            const int h = D.trou[o + t], q = D.part[o + t];
            ok = h >= 1 && h <= n_so && q >= 1 && q <= n_so;
        }
        if (!ok) atomicAdd(n_bad, 1ull);
    }
}

// cip/hdig.f:40-80, fdiag(s) = F(iorb(s), iorb(s)) as cip/ijkf.f:33-36 sets it.
__global__ void k_hdig(DetArrays D, int n_det, const double* __restrict__ fock,
                       const double* __restrict__ jint, const double* __restrict__ kint,
                       double* __restrict__ out) {
#pragma clang fp contract(off)
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n_det) return;
    const int nei = D.ne[d];
    double ab = 0.0;
    const int l1 = D.nd[d];
    for (int l = l1; l < l1 + nei; ++l) {
        const int s1p = D.part[l], s1t = D.trou[l];
        const int n1p = D.iorb[s1p - 1], n1t = D.iorb[s1t - 1];
        ab = ab + fock[tri1(n1p, n1p)] - fock[tri1(n1t, n1t)];
        const int ns1t = D.ispin[s1t - 1], ns1p = D.ispin[s1p - 1];
        for (int j = l1; j <= l; ++j) {
            const int s2t = D.trou[j], s2p = D.part[j];
            const int ns2t = D.ispin[s2t - 1], ns2p = D.ispin[s2p - 1];
            const int n2t = D.iorb[s2t - 1], n2p = D.iorb[s2p - 1];
            const int64_t n12t = tri1s(n1t, n2t), n12p = tri1s(n1p, n2p);
            const int64_t n1tp = tri1(n2p, n1t), n2tp = tri1(n1p, n2t);
            double ca = jint[n12t] + jint[n12p] - jint[n1tp] - jint[n2tp];
            if (ns1t == ns2t) ca = ca - kint[n12t];
            if (ns1p == ns2p) ca = ca - kint[n12p];
            if (ns1t == ns2p) ca = ca + kint[n1tp];
            if (ns1p == ns2t) ca = ca + kint[n2tp];
            if (j == l) ca = 0.5 * ca;
            ab = ab + ca;
        }
    }
    out[d] = ab;
}

// cip/hntd.f:36-168 for one pair (da, db): the determinant with more excitations is
// "1"; its excitation list is reordered to line up with the other's (each swap flips the
// sign); unmatched spin-orbitals give nv1 -> nv2; one difference: F(n1, n2) corrected
// for the holes and particles of determinant 1; two: (n1 n2|n3 n4) - (n1 n4|n3 n2) by
// spin; identical determinants or more than two differences: 0.
__device__ double hntd_pair(const DetArrays& D, int da, int db, int norb,
                            const double* __restrict__ fock, const double* __restrict__ mo) {
#pragma clang fp contract(off)
    int ita[2 * kMaxExc], itb[2 * kMaxExc];
    int nv1[2] = {0, 0}, nv2[2] = {0, 0}, natu1[2] = {0, 0}, natu2[2] = {0, 0};
    const int kr = (D.ne[da] - D.ne[db] <= 0) ? db : da;
    const int kl = (kr == db) ? da : db;
    const int ne1 = D.ne[kr];
    if (ne1 == 0) return 0.0;
    const int ne2 = D.ne[kl];
    int nbdif = ne1 - ne2;
    if (nbdif > 2) return 0.0;
    for (int t = 0; t < ne1; ++t) {
        ita[t] = D.trou[D.nd[kr] + t];
        ita[t + ne1] = D.part[D.nd[kr] + t];
    }
    bool ysig = false;
    if (ne2 > 0) {
        for (int t = 0; t < ne2; ++t) {
            itb[t] = D.trou[D.nd[kl] + t];
            itb[t + ne2] = D.part[D.nd[kl] + t];
        }
        for (int k = 0; k < 2; ++k) {               // holes, then particles
            const int k1 = k * ne1, k2 = k * ne2;
            for (int i = 0; i < ne2; ++i) {
                const int nit1 = itb[k2 + i];
                bool found = false;
                for (int j = k1; j < k1 + ne1; ++j) {
                    if (ita[j] != nit1) continue;
                    if (i != j - k1) {
                        ita[j] = ita[i + k1];
                        ita[i + k1] = nit1;
                        ysig = !ysig;
                    }
                    found = true;
                    break;
                }
                if (!found) {
                    ++nbdif;
                    if (nbdif > 2) return 0.0;
                    nv1[nbdif - 1] = nit1;
                    natu1[nbdif - 1] = k + 1;
                    natu2[nbdif - 1] = i;
                }
            }
        }
    }
    if (nbdif <= 0) return 0.0;
    const int nar = ne1 - ne2;
    for (int i = 0; i < nar; ++i) {
        nv1[i] = ita[ne1 - 1 - i];
        nv2[i] = ita[2 * ne1 - 1 - i];
    }
    for (int i = nar; i < nbdif; ++i) {
        const int k = natu1[i], ni = natu2[i], nu = nv1[i];
        nv2[i] = ita[ni + ne1 * (k - 1)];
        if (k == 1) {
            nv1[i] = nv2[i];
            nv2[i] = nu;
            ysig = !ysig;
        }
    }
    const int n1 = D.iorb[nv1[0] - 1], n2 = D.iorb[nv2[0] - 1];
    const int ns1 = D.ispin[nv1[0] - 1], ns2 = D.ispin[nv2[0] - 1];
    double ha = 0.0;
    if (nbdif == 2) {
        const int n3 = D.iorb[nv1[1] - 1], n4 = D.iorb[nv2[1] - 1];
        const int ns4 = D.ispin[nv2[1] - 1];
        if (ns1 == ns2) ha = ai1(mo, n1, n2, n3, n4);
        if (ns1 == ns4) ha = ha - ai1(mo, n1, n4, n3, n2);
    } else {
        ha = fock[tri1s(n1, n2)];
        for (int i = 0; i < ne1; ++i) {
            int np = D.iorb[ita[i] - 1];
            int nsu = D.ispin[ita[i] - 1];
            ha = ha - ai1(mo, n1, n2, np, np);
            if (ns1 == nsu) ha = ha + ai1(mo, n1, np, n2, np);
            np = D.iorb[ita[i + ne1] - 1];
            if (np > norb) continue;
            nsu = D.ispin[ita[i + ne1] - 1];
            ha = ha + ai1(mo, n1, n2, np, np);
            if (ns1 == nsu) ha = ha - ai1(mo, n1, np, n2, np);
        }
    }
    return ysig ? -ha : ha;
}

__global__ void k_hntd(DetArrays D, int norb, const int* __restrict__ pairs, int64_t n_pairs,
                       int n_det, const double* __restrict__ fock,
                       const double* __restrict__ mo, double* __restrict__ out,
                       unsigned long long* __restrict__ n_bad) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_pairs; q += stride) {
        const int da = pairs[2 * q], db = pairs[2 * q + 1];
        if (da < 0 || db < 0 || da >= n_det || db >= n_det) {
            out[q] = __builtin_nan("");
            atomicAdd(n_bad, 1ull);
            continue;
        }
        out[q] = hntd_pair(D, da, db, norb, fock, mo);
    }
}


int sc_validate(cipsi_ijkl_ctx* ctx, const DetArrays& D, int n_det, int64_t n_exc, int n_mo,
                unsigned long long* n_bad) {
    if (hipMemsetAsync(n_bad, 0, sizeof(unsigned long long), ctx->stream) != hipSuccess)
        return CIPSI_ERR_HIP;
    const int64_t work = n_det > 2 * n_mo ? n_det : 2 * n_mo;
    const unsigned grid = (unsigned)((work + 255) / 256 < 65536 ? (work + 255) / 256 : 65536);
    hipLaunchKernelGGL(k_sc_validate, dim3(grid), dim3(256), 0, ctx->stream, D, n_det, n_exc,
                       2 * n_mo, n_mo, n_bad);
    return hipGetLastError() == hipSuccess ? CIPSI_OK : CIPSI_ERR_HIP;
}

int sc_finish(cipsi_ijkl_ctx* ctx, unsigned long long* n_bad, const char* what) {
    unsigned long long bad = 0;
    if (hipMemcpyAsync(&bad, n_bad, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return CIPSI_ERR_HIP;
    if (bad) {
        ctx->last_error = std::string(what) + ": " + std::to_string(bad) +
                          " determinants, spin-orbitals or pairs out of range";
        return CIPSI_ERR_ARG;
    }
    return CIPSI_OK;
}

}  // namespace

extern "C" {

int cipsi_ijkl_gather_dev(cipsi_ijkl_ctx* ctx, int n_mo, const double* d_mo_packed,
                          const int32_t* d_labels, int64_t n, double* d_out) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if (n_mo < 1 || n < 0 || !d_mo_packed || (n > 0 && (!d_labels || !d_out)))
        return CIPSI_ERR_ARG;
    if (n_mo > 255) return CIPSI_ERR_LABEL;
    return gather_impl(ctx, n_mo, d_mo_packed, d_labels, n, d_out);
}

int cipsi_ijkl_fock_dev(cipsi_ijkl_ctx* ctx, int n_mo, int noc, const double* d_mo_packed,
                        const double* d_h_mo_packed, double* d_energy, double* d_fock,
                        double* d_j, double* d_k) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_fock_args(n_mo, noc))) return rc;
    if (!d_mo_packed || !d_h_mo_packed || !d_energy || !d_fock || !d_j || !d_k)
        return CIPSI_ERR_ARG;
    return fock_impl(ctx, n_mo, noc, d_mo_packed, d_h_mo_packed, d_energy, d_fock, d_j, d_k);
}

int cipsi_ijkl_fock(cipsi_ijkl_ctx* ctx, int n_mo, int noc, const double* mo_packed,
                    const double* h_mo_packed, double* energy, double* fock, double* jint,
                    double* kint) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if ((rc = check_fock_args(n_mo, noc))) return rc;
    if (!mo_packed || !h_mo_packed || !energy || !fock || !jint || !kint) return CIPSI_ERR_ARG;
    const int64_t pm = pair0(n_mo, 0);
    const size_t s_mo = sizeof(double) * (size_t)pair0(pm, 0), s_p = sizeof(double) * pm;
    if ((rc = ctx_alloc(ctx, ctx->io_in, s_mo))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->io_in2, s_p))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->io_out, 3 * s_p))) return rc;
    unsigned long long* n_bad;
    double* d_e;
    if ((rc = cons_scalars(ctx, &n_bad, &d_e))) return rc;
    double* d_f = (double*)ctx->io_out.ptr;
    double* d_j = d_f + pm;
    double* d_k = d_j + pm;
    hipStream_t st = ctx->stream;
    if (hipMemcpyAsync(ctx->io_in.ptr, mo_packed, s_mo, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(ctx->io_in2.ptr, h_mo_packed, s_p, hipMemcpyHostToDevice, st) != hipSuccess)
        return CIPSI_ERR_HIP;
    if ((rc = fock_impl(ctx, n_mo, noc, (const double*)ctx->io_in.ptr,
                        (const double*)ctx->io_in2.ptr, d_e, d_f, d_j, d_k)))
        return rc;
    if (hipMemcpyAsync(energy, d_e, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(fock, d_f, s_p, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(jint, d_j, s_p, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(kint, d_k, s_p, hipMemcpyDeviceToHost, st) != hipSuccess)
        return CIPSI_ERR_HIP;
    return hipStreamSynchronize(st) == hipSuccess ? CIPSI_OK : CIPSI_ERR_HIP;
}


int cipsi_ijkl_hdig_dev(cipsi_ijkl_ctx* ctx, int n_mo, int n_det, int64_t n_exc,
                        const int* d_ne, const int* d_nd, const int* d_trou, const int* d_part,
                        const int* d_iorb, const int* d_ispin, const double* d_fock,
                        const double* d_jint, const double* d_kint, double* d_out) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if (n_mo < 1 || n_det < 0 || n_exc < 0 || !d_iorb || !d_ispin || !d_fock || !d_jint ||
        !d_kint || (n_det > 0 && (!d_ne || !d_nd || !d_out)) ||
        (n_exc > 0 && (!d_trou || !d_part)))
        return CIPSI_ERR_ARG;
    if (n_mo > kMaxMo) return CIPSI_ERR_UNSUPPORTED;
    unsigned long long* n_bad;
    double* e;
    if ((rc = cons_scalars(ctx, &n_bad, &e))) return rc;
    const DetArrays D{d_ne, d_nd, d_trou, d_part, d_iorb, d_ispin};
    if ((rc = sc_validate(ctx, D, n_det, n_exc, n_mo, n_bad))) return rc;
    if ((rc = sc_finish(ctx, n_bad, "hdig"))) return rc;
    if (n_det > 0) {
        StageTimer tm(ctx, "hdig");
        hipLaunchKernelGGL(k_hdig, dim3((n_det + 255) / 256), dim3(256), 0, ctx->stream, D, n_det,
                           d_fock, d_jint, d_kint, d_out);
        if (hipGetLastError() != hipSuccess) return CIPSI_ERR_HIP;
    }
    return hipStreamSynchronize(ctx->stream) == hipSuccess ? CIPSI_OK : CIPSI_ERR_HIP;
}

int cipsi_ijkl_hntd_dev(cipsi_ijkl_ctx* ctx, int n_mo, int n_det, int64_t n_exc,
                        const int* d_ne, const int* d_nd, const int* d_trou, const int* d_part,
                        const int* d_iorb, const int* d_ispin, const int* d_pairs,
                        int64_t n_pairs, const double* d_mo_packed, const double* d_fock,
                        double* d_out) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if (n_mo < 1 || n_det < 0 || n_exc < 0 || n_pairs < 0 || !d_iorb || !d_ispin || !d_fock ||
        !d_mo_packed || (n_det > 0 && (!d_ne || !d_nd)) ||
        (n_exc > 0 && (!d_trou || !d_part)) || (n_pairs > 0 && (!d_pairs || !d_out)))
        return CIPSI_ERR_ARG;
    if (n_mo > kMaxMo) return CIPSI_ERR_UNSUPPORTED;
    unsigned long long* n_bad;
    double* e;
    if ((rc = cons_scalars(ctx, &n_bad, &e))) return rc;
    const DetArrays D{d_ne, d_nd, d_trou, d_part, d_iorb, d_ispin};
    if ((rc = sc_validate(ctx, D, n_det, n_exc, n_mo, n_bad))) return rc;
    if ((rc = sc_finish(ctx, n_bad, "hntd"))) return rc;
    if (n_pairs > 0) {
        StageTimer tm(ctx, "hntd");
        const int64_t blocks = (n_pairs + 255) / 256;
        hipLaunchKernelGGL(k_hntd, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256), 0,
                           ctx->stream, D, n_mo, d_pairs, n_pairs, n_det, d_fock, d_mo_packed,
                           d_out, n_bad);
        if (hipGetLastError() != hipSuccess) return CIPSI_ERR_HIP;
    }
    return sc_finish(ctx, n_bad, "hntd");
}

}  // extern "C"
