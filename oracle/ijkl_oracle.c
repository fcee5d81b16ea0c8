/*
 * ijkl_oracle.c -- CPU restatement of the reference ijkl AO->MO transform (C1 path).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the host-core
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and only as the checker / the timed CPU baseline; the product path
 * (cipsi_amd) never calls it.
 *
 * It restates, in C, the loop nests and the summation order of the reference
 * Fortran routines for the C1 (single symmetry block, nt = 1) case:
 *
 *   or_transd_c1   CIPSI_sources/ijkl/transd.f:85-97,210-218,356-373
 *                  (label decode iword_func.f:2-52; producer pshf/qout.f:71-84)
 *   or_tpqkld_c1   CIPSI_sources/ijkl/tpqkld.f:25-104  (first half, Q1 + Q2)
 *   or_tijkld_c1   CIPSI_sources/ijkl/tijkld.f:28-88   (second half, Q3 + Q4)
 *   or_tmono       CIPSI_sources/ijkl/tmono.f:39-54    (one-electron transform)
 *   or_fock        CIPSI_sources/fock/fock.f:80-127    (closed-shell energy, Fock, J, K)
 *   or_hdig        CIPSI_sources/cip/hdig.f:40-80       (determinant energy)
 *   or_hntd        CIPSI_sources/cip/hntd.f:36-168      (off-diagonal Slater-Condon element)
 *
 * Every accumulation starts from 0.0 and adds in the same index order as the
 * Fortran loop; build with -ffp-contract=off so that no multiply-add is fused.
 * The results are then bit-identical to a flang build of the reference without FMA.
 *
 * Layouts (all 0-based here; the reference is 1-based):
 *   pair0(a,b) = a(a+1)/2 + b           for a >= b       (itijkl.f:38, transd.f:32)
 *   AO packed supermatrix  V[pair0(ij, kl)], ij >= kl    (transd.f:369-371)
 *   C  column-major N x M, C(mu, m) = c[mu + m*ldc]       (vector.f:38-41)
 *   half-transformed  H[npq * P_M + ilk], ilk = pair0(l, k), l >= k  (tpqkld.f:91-102)
 *   MO packed  out[pair0(ilk, nij)], nij = pair0(j, i), nij <= ilk   (tijkld.f:31-79)
 *
 * The *_range variants compute a contiguous sub-range of the outer loop with
 * exactly the same per-element arithmetic, so a sampled or OpenMP-parallel run
 * produces the same bits as the serial one.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

static inline int64_t pair0(int64_t a, int64_t b) { return a * (a + 1) / 2 + b; }

/* Label codec, iword_func.f:2-52 (decode) and pshf/iword_func.f:2-14 (encode). */
int32_t or_label_encode(int i, int j, int k, int l)
{
    return (int32_t)(((uint32_t)i << 24) | ((uint32_t)j << 16) | ((uint32_t)k << 8) | (uint32_t)l);
}

void or_label_decode(int32_t label, int *i, int *j, int *k, int *l)
{
    uint32_t u = (uint32_t)label;
    *i = (int)((u >> 24) & 255u);
    *j = (int)((u >> 16) & 255u);
    *k = (int)((u >> 8) & 255u);
    *l = (int)(u & 255u);
}

/*
 * transd, C1 route (nt = 1, one symmetry block, icas = 8 -> label 1000/1010):
 * the block is zeroed (itijkl.f:191-192), then every label up to the first zero
 * label (transd.f:93) is decoded, its value rescaled by 2 for i==j, for k==l and
 * for (i,j)==(k,l) (transd.f:210-218), and stored at
 *   ij1 = i(i-1)/2 + j, kl1 = k(k-1)/2 + l, ijkl = ij1(ij1-1)/2 + kl1  (1-based)
 * (transd.f:369-372).  Later duplicates overwrite earlier ones.
 *
 * Returns the number of labels consumed, or -(1 + index) of the first label whose
 * target position falls outside the block (the reference would write out of bounds).
 */
int64_t or_transd_c1(int n_ao, const double *xx, const int32_t *ix, int64_t n_entries,
                     double *bijkl)
{
    const int64_t pa = (int64_t)n_ao * (n_ao + 1) / 2;
    const int64_t nblk = pa * (pa + 1) / 2;
    memset(bijkl, 0, (size_t)nblk * sizeof(double));
    int64_t m;
    for (m = 0; m < n_entries; ++m) {
        int32_t lab = ix[m];
        if (lab == 0) break;
        int i, j, k, l;
        or_label_decode(lab, &i, &j, &k, &l);
        double val = xx[m];                 /* qq4 = 1 for nt = 1 */
        if (i == j) val = val + val;
        if (k == l) val = val + val;
        if (i == k && j == l) val = val + val;
        if (i < 1 || j < 1 || k < 1 || l < 1 || i > n_ao || j > n_ao || k > n_ao || l > n_ao)
            return -(1 + m);
        int64_t ij1 = (int64_t)i * (i - 1) / 2 + j;
        int64_t kl1 = (int64_t)k * (k - 1) / 2 + l;
        int64_t ijkl = ij1 * (ij1 - 1) / 2 + kl1;       /* 1-based */
        if (ijkl < 1 || ijkl > nblk) return -(1 + m);
        bijkl[ijkl - 1] = val;
    }
    return m;
}

/*
 * tpqkld (tpqkld.f:25-104) for the C1 block, all symmetry flags true.
 * For each AO pair npq (p >= q, npq ascending):
 *   tab(r,s) = V[pair(npq, nrs)] mirrored to a full symmetric N x N (57-80);
 *   Q1: tac(r) = sum_{s=1..N} c(s,l) * tab(r,s)         for every l   (82-90)
 *   Q2: H(npq, (l,k)) = sum_{r=1..N} c(r,k) * tac(r)     for k <= l    (91-102)
 * Output is pq-major: H[npq*P_M + pair0(l,k)].
 */
void or_tpqkld_c1_range(int n_ao, int n_mo, const double *bijkl, const double *c, int ldc,
                        double *h, int64_t npq_begin, int64_t npq_end)
{
    const int n = n_ao;
    const int64_t pm = (int64_t)n_mo * (n_mo + 1) / 2;
    double *tab = (double *)malloc(sizeof(double) * (size_t)n * n);
    double *tac = (double *)malloc(sizeof(double) * (size_t)n);
    for (int p = 0; p < n; ++p) {
        for (int q = 0; q <= p; ++q) {
            const int64_t npq = pair0(p, q);
            if (npq < npq_begin || npq >= npq_end) continue;
            /* gather + mirror: tab[r*n + s] = tab[s*n + r] = (pq|rs) */
            for (int r = 0; r < n; ++r) {
                for (int s = 0; s <= r; ++s) {
                    const int64_t nrs = pair0(r, s);
                    double v = (npq >= nrs) ? bijkl[pair0(npq, nrs)] : bijkl[pair0(nrs, npq)];
                    tab[(size_t)r * n + s] = v;
                    tab[(size_t)s * n + r] = v;
                }
            }
            double *hrow = h + npq * pm;
            int64_t mono = 0;
            for (int l = 0; l < n_mo; ++l) {
                const double *cl = c + (size_t)l * ldc;
                for (int r = 0; r < n; ++r) {
                    double sum = 0.0;
                    const double *tr = tab + (size_t)r * n;
                    for (int s = 0; s < n; ++s) sum = sum + cl[s] * tr[s];
                    tac[r] = sum;
                }
                for (int k = 0; k <= l; ++k) {
                    const double *ck = c + (size_t)k * ldc;
                    double sum = 0.0;
                    for (int r = 0; r < n; ++r) sum = sum + ck[r] * tac[r];
                    hrow[mono++] = sum;
                }
            }
        }
    }
    free(tab);
    free(tac);
}

void or_tpqkld_c1(int n_ao, int n_mo, const double *bijkl, const double *c, int ldc, double *h)
{
    const int64_t pa = (int64_t)n_ao * (n_ao + 1) / 2;
    or_tpqkld_c1_range(n_ao, n_mo, bijkl, c, ldc, h, 0, pa);
}

/*
 * tijkld (tijkld.f:28-88) for the C1 block.
 * For each MO pair ilk = (l >= k), ascending:
 *   tab(p,q) = H(pair(p,q), ilk) gathered with stride P_M, mirrored (35-49);
 *   Q3: tac(p) = sum_{q=1..N} c(q,j) tab(p,q)      for j = 1, 2, ...   (53-60)
 *   Q4: out(ilk, (j,i)) = sum_{p=1..N} c(p,i) tac(p), i <= j,
 *       stopping as soon as nij == ilk (yijkl, 61-76) -- so Q3 runs only up to j = l.
 * Output: out[pair0(ilk, nij)], the packed MO stream of tijkld.f:68-74.
 */
void or_tijkld_c1_range(int n_ao, int n_mo, const double *h, const double *c, int ldc,
                        double *out, int64_t ilk_begin, int64_t ilk_end)
{
    const int n = n_ao;
    const int64_t pm = (int64_t)n_mo * (n_mo + 1) / 2;
    double *tab = (double *)malloc(sizeof(double) * (size_t)n * n);
    double *tac = (double *)malloc(sizeof(double) * (size_t)n);
    for (int l = 0; l < n_mo; ++l) {
        for (int k = 0; k <= l; ++k) {
            const int64_t ilk = pair0(l, k);
            if (ilk < ilk_begin || ilk >= ilk_end) continue;
            for (int p = 0; p < n; ++p) {
                for (int q = 0; q <= p; ++q) {
                    double v = h[pair0(p, q) * pm + ilk];
                    tab[(size_t)p * n + q] = v;
                    tab[(size_t)q * n + p] = v;
                }
            }
            double *orow = out + pair0(ilk, 0);
            int64_t nij = 0;
            int done = 0;
            for (int j = 0; j < n_mo && !done; ++j) {
                const double *cj = c + (size_t)j * ldc;
                for (int p = 0; p < n; ++p) {
                    double sum = 0.0;
                    const double *tp = tab + (size_t)p * n;
                    for (int q = 0; q < n; ++q) sum = sum + cj[q] * tp[q];
                    tac[p] = sum;
                }
                for (int i = 0; i <= j; ++i) {
                    const double *ci = c + (size_t)i * ldc;
                    double sum = 0.0;
                    for (int p = 0; p < n; ++p) sum = sum + ci[p] * tac[p];
                    orow[nij] = sum;
                    if (nij == ilk) { done = 1; break; }
                    ++nij;
                }
            }
        }
    }
    free(tab);
    free(tac);
}

void or_tijkld_c1(int n_ao, int n_mo, const double *h, const double *c, int ldc, double *out)
{
    const int64_t pm = (int64_t)n_mo * (n_mo + 1) / 2;
    or_tijkld_c1_range(n_ao, n_mo, h, c, ldc, out, 0, pm);
}

/* Whole C1 transform from the packed AO block: tpqkld then tijkld (itijkl.f:214-285). */
int or_ijkl_c1(int n_ao, int n_mo, const double *bijkl, const double *c, int ldc, double *out)
{
    const int64_t pa = (int64_t)n_ao * (n_ao + 1) / 2;
    const int64_t pm = (int64_t)n_mo * (n_mo + 1) / 2;
    double *h = (double *)malloc(sizeof(double) * (size_t)(pa * pm));
    if (!h) return -1;
    or_tpqkld_c1(n_ao, n_mo, bijkl, c, ldc, h);
    or_tijkld_c1(n_ao, n_mo, h, c, ldc, out);
    free(h);
    return 0;
}

/*
 * OpenMP variants: outer loops split over threads; each output element is
 * produced by the same sequential inner loops, so results equal the serial ones.
 */
#ifdef _OPENMP
#include <omp.h>
#endif
void or_tpqkld_c1_omp(int n_ao, int n_mo, const double *bijkl, const double *c, int ldc,
                      double *h, int64_t npq_begin, int64_t npq_end, int nthreads)
{
#ifdef _OPENMP
    const int64_t span = npq_end - npq_begin;
#pragma omp parallel num_threads(nthreads)
    {
        int t = omp_get_thread_num(), nt = omp_get_num_threads();
        int64_t b = npq_begin + span * t / nt, e = npq_begin + span * (t + 1) / nt;
        or_tpqkld_c1_range(n_ao, n_mo, bijkl, c, ldc, h, b, e);
    }
#else
    (void)nthreads;
    or_tpqkld_c1_range(n_ao, n_mo, bijkl, c, ldc, h, npq_begin, npq_end);
#endif
}

void or_tijkld_c1_omp_list(int n_ao, int n_mo, const double *h, const double *c, int ldc,
                           double *out, const int64_t *ilk_list, int64_t n_list, int nthreads)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int64_t t = 0; t < n_list; ++t)
        or_tijkld_c1_range(n_ao, n_mo, h, c, ldc, out, ilk_list[t], ilk_list[t] + 1);
#ifndef _OPENMP
    (void)nthreads;
#endif
}

void or_tpqkld_c1_omp_list(int n_ao, int n_mo, const double *bijkl, const double *c, int ldc,
                           double *h, const int64_t *npq_list, int64_t n_list, int nthreads)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int64_t t = 0; t < n_list; ++t)
        or_tpqkld_c1_range(n_ao, n_mo, bijkl, c, ldc, h, npq_list[t], npq_list[t] + 1);
#ifndef _OPENMP
    (void)nthreads;
#endif
}

/*
 * tmono (tmono.f:39-54):  (hC)(j,i) = sum_k C(k,i) h(k,j)  over k = 1..N, then
 * hmo(i,j) = sum_k (hC)(k,i) C(k,j) for j <= i, packed i-outer / j-inner.
 */
void or_tmono(int n_ao, int n_mo, const double *hpk, const double *c, int ldc, double *hmo)
{
    const int n = n_ao;
    double *hc = (double *)malloc(sizeof(double) * (size_t)n * n_mo);
    for (int i = 0; i < n_mo; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) {
                int64_t kj = (k >= j) ? pair0(k, j) : pair0(j, k);
                s = s + c[(size_t)i * ldc + k] * hpk[kj];
            }
            hc[(size_t)i * n + j] = s;
        }
    int64_t ij = 0;
    for (int i = 0; i < n_mo; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s = s + hc[(size_t)i * n + k] * c[(size_t)j * ldc + k];
            hmo[ij++] = s;
        }
    free(hc);
}

/*
 * Energy and gradient oracle for E(C) = sum_{pq} h^MO_pq gamma_pq
 *                                     + 1/2 sum_{pqrs} Gamma_pqrs (pq|rs)_MO.
 * gamma is a full symmetric M x M matrix (row-major); Gamma is given on the MO
 * pair supermatrix, packed like the MO integrals (Gamma[pair0(a,b)], a >= b,
 * a = pair0(p,q)), and Gamma_pqrs = Gamma[pair(p,q), pair(r,s)] for all orders.
 * The full 4-index sum equals sum over packed pairs a,b of m(a) m(b) G_ab g_ab
 * with m(a) = 2 for p != q, 1 for p == q.
 *
 * Gradient:  dE1/dC = 2 h C gamma;
 *            dE2/dC(mu,p) = 2 sum_{q,nu} C(nu,q) L[pair(mu,nu)][pair(p,q)],
 *            L[mn][a] = sum_b H[mn][b] m(b) Gamma[a][b],  H = first half (tpqkld).
 * Serial FP64 loops; grad is column-major N x M with leading dimension N.
 */
static inline int64_t pairsym(int64_t a, int64_t b) { return a >= b ? pair0(a, b) : pair0(b, a); }

double or_energy(int n_ao, int n_mo, const double *mo_packed, const double *hmo_packed,
                 const double *gamma, const double *Gamma_packed)
{
    const int64_t pm = (int64_t)n_mo * (n_mo + 1) / 2;
    double *mult = (double *)malloc(sizeof(double) * (size_t)pm);
    for (int p = 0; p < n_mo; ++p)
        for (int q = 0; q <= p; ++q) mult[pair0(p, q)] = (p == q) ? 1.0 : 2.0;
    double e1 = 0.0;
    for (int p = 0; p < n_mo; ++p)
        for (int q = 0; q < n_mo; ++q)
            e1 += hmo_packed[pairsym(p, q)] * gamma[(size_t)p * n_mo + q];
    double e2 = 0.0;
    int64_t idx = 0;
    for (int64_t a = 0; a < pm; ++a)
        for (int64_t b = 0; b <= a; ++b, ++idx) {
            const double w = (a == b) ? 1.0 : 2.0;
            e2 += w * mult[a] * mult[b] * Gamma_packed[idx] * mo_packed[idx];
        }
    free(mult);
    (void)n_ao;
    return e1 + 0.5 * e2;
}

void or_energy_grad(int n_ao, int n_mo, const double *bijkl, const double *hpk,
                    const double *c, int ldc, const double *gamma, const double *Gamma_packed,
                    double *energy, double *grad)
{
    const int n = n_ao, m = n_mo;
    const int64_t pa = (int64_t)n * (n + 1) / 2;
    const int64_t pm = (int64_t)m * (m + 1) / 2;
    double *h = (double *)malloc(sizeof(double) * (size_t)(pa * pm));
    double *out = (double *)malloc(sizeof(double) * (size_t)(pm * (pm + 1) / 2));
    double *hmo = (double *)malloc(sizeof(double) * (size_t)pm);
    or_tpqkld_c1(n, m, bijkl, c, ldc, h);
    or_tijkld_c1(n, m, h, c, ldc, out);
    or_tmono(n, m, hpk, c, ldc, hmo);
    *energy = or_energy(n, m, out, hmo, gamma, Gamma_packed);

    /* multiplicities of MO pairs */
    double *mult = (double *)malloc(sizeof(double) * (size_t)pm);
    for (int p = 0; p < m; ++p)
        for (int q = 0; q <= p; ++q) mult[pair0(p, q)] = (p == q) ? 1.0 : 2.0;
    /* L = H * Gw, Gw[b][a] = m(b) Gamma[a][b] */
    double *L = (double *)calloc((size_t)(pa * pm), sizeof(double));
    for (int64_t mn = 0; mn < pa; ++mn) {
        const double *hr = h + mn * pm;
        double *lr = L + mn * pm;
        for (int64_t a = 0; a < pm; ++a) {
            double s = 0.0;
            for (int64_t b = 0; b < pm; ++b) s += hr[b] * mult[b] * Gamma_packed[pairsym(a, b)];
            lr[a] = s;
        }
    }
    /* dE2/dC(mu,p) = 2 sum_{nu,q} C(nu,q) L[pair(mu,nu)][pair(p,q)]  (all nu, q) */
    for (int p = 0; p < m; ++p)
        for (int mu = 0; mu < n; ++mu) {
            double s = 0.0;
            for (int nu = 0; nu < n; ++nu) {
                const double *lr = L + pairsym(mu, nu) * pm;
                for (int q = 0; q < m; ++q) s += c[(size_t)q * ldc + nu] * lr[pairsym(p, q)];
            }
            grad[(size_t)p * n + mu] = 2.0 * s;
        }
    /* dE1/dC = 2 h C gamma */
    for (int p = 0; p < m; ++p)
        for (int mu = 0; mu < n; ++mu) {
            double s = 0.0;
            for (int q = 0; q < m; ++q) {
                double hc = 0.0;
                for (int nu = 0; nu < n; ++nu) hc += hpk[pairsym(mu, nu)] * c[(size_t)q * ldc + nu];
                s += hc * gamma[(size_t)q * m + p];
            }
            grad[(size_t)p * n + mu] += 2.0 * s;
        }
    free(L);
    free(mult);
    free(h);
    free(out);
    free(hmo);
}

/*
 * Integral consumer of the packed MO stream: the closed-shell reference energy and the
 * MO Fock operator of CIPSI_sources/fock/fock.f:80-127, with the integrals fetched as
 * ai() fetches them for C1 (fock/ai.f:12-35: pair indices, larger pair first).
 * noc doubly occupied orbitals; hmo packed M(M+1)/2 (h_MO, i >= j); outputs packed:
 *   fock[pair0(i,j)]  F_ij = h_ij + sum_{k<noc} 2 (ij|kk) - (ik|jk), then |F| < 1e-7 -> 0
 *   jint[pair0(i,j)]  (ii|jj),  kint[pair0(i,j)]  (ij|ij)        (fock.f:108-112)
 * The energy is val of fock.f:80-100 (without enuc).  Loop and summation order follow
 * fock.f so the results match a build of the reference without FMA.
 */
static inline double or_ai(const double *mo, int i, int j, int k, int l)
{
    const int64_t ij = (i >= j) ? pair0(i, j) : pair0(j, i);
    const int64_t kl = (k >= l) ? pair0(k, l) : pair0(l, k);
    return (ij >= kl) ? mo[pair0(ij, kl)] : mo[pair0(kl, ij)];
}

double or_fock(int n_mo, int noc, const double *mo, const double *hmo, double *fock,
               double *jint, double *kint)
{
    const int m = n_mo;
    double *f = (double *)malloc(sizeof(double) * (size_t)m * m);   /* f(i,j), i >= j used */
    for (int i = 0; i < m; ++i)
        for (int j = 0; j <= i; ++j) f[(size_t)i * m + j] = hmo[pair0(i, j)];
    double val = 0.0;
    for (int i = 0; i < noc; ++i) {                                /* fock.f:81-99 */
        val = val + f[(size_t)i * m + i] * 2.0;
        double tt = or_ai(mo, i, i, i, i);
        f[(size_t)i * m + i] = f[(size_t)i * m + i] + tt;
        val = val + tt;
        for (int k = 0; k < i; ++k) {
            tt = or_ai(mo, i, i, k, k);
            const double tk = or_ai(mo, i, k, i, k);
            tt = tt + tt - tk;
            f[(size_t)i * m + i] = f[(size_t)i * m + i] + tt;
            f[(size_t)k * m + k] = f[(size_t)k * m + k] + tt;
            val = val + tt + tt;
            for (int j = 0; j < noc; ++j)
                f[(size_t)i * m + k] = f[(size_t)i * m + k]
                                       + (or_ai(mo, i, k, j, j) * 2.0 - or_ai(mo, i, j, k, j));
        }
    }
    for (int i = noc; i < m; ++i)                                  /* fock.f:102-109 */
        for (int j = 0; j <= i; ++j) {
            double tt = 0.0;
            for (int k = 0; k < noc; ++k)
                tt = tt + (2.0 * or_ai(mo, i, j, k, k) - or_ai(mo, i, k, j, k));
            f[(size_t)i * m + j] = f[(size_t)i * m + j] + tt;
        }
    for (int i = 0; i < m; ++i)                                    /* fock.f:110-114, 124-128 */
        for (int j = 0; j <= i; ++j) {
            jint[pair0(i, j)] = or_ai(mo, i, i, j, j);
            kint[pair0(i, j)] = or_ai(mo, i, j, i, j);
            const double v = f[(size_t)i * m + j];
            fock[pair0(i, j)] = (fabs(v) < 1.0e-7) ? 0.0 : v;
        }
    free(f);
    return val;
}

/*
 * Slater-Condon elements of cip between determinants held as excitations of the
 * closed-shell reference (cip's /spo/ arrays, all numbers 1-based as there):
 *   determinant d has ne[d] excitations: holes trou[nd[d] + t] and particles
 *   part[nd[d] + t], t < ne[d], as spin-orbital numbers; spin-orbital s has orbital
 *   iorb[s - 1] and spin ispin[s - 1].
 *   fock, jint, kint: packed F, J = (ii|jj), K = (ij|ij) (or_fock, read by cip in
 *   ijkf.f:26-36); mo: the packed MO stream, fetched as ai() fetches it.
 * tri1(a, b) = a(a-1)/2 + b - 1 is the 0-based packed index of the 1-based pair (a, b)
 * as cip's num(a) + b, without reordering (hdig.f uses it as written).
 */
static inline int64_t tri1(int64_t a, int64_t b) { return a * (a - 1) / 2 + b - 1; }
static inline int64_t tri1s(int64_t a, int64_t b) { return a >= b ? tri1(a, b) : tri1(b, a); }

/* hdig.f:40-80: E(D) - E(reference); fdiag(s) = F(iorb(s), iorb(s)) (ijkf.f:33-36). */
double or_hdig(int d, const int *ne, const int *nd, const int *trou, const int *part,
               const int *iorb, const int *ispin, const double *fock, const double *jint,
               const double *kint)
{
    const int nei = ne[d];
    if (nei == 0) return 0.0;
    const int l1 = nd[d];
    double ab = 0.0;
    for (int l = l1; l < l1 + nei; ++l) {
        const int s1p = part[l], s1t = trou[l];
        const int n1p = iorb[s1p - 1], n1t = iorb[s1t - 1];
        ab = ab + fock[tri1(n1p, n1p)] - fock[tri1(n1t, n1t)];
        const int ns1t = ispin[s1t - 1], ns1p = ispin[s1p - 1];
        for (int j = l1; j <= l; ++j) {
            const int s2t = trou[j], s2p = part[j];
            const int ns2t = ispin[s2t - 1], ns2p = ispin[s2p - 1];
            const int n2t = iorb[s2t - 1], n2p = iorb[s2p - 1];
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
    return ab;
}

/* ai() of cip for C1 on the packed MO stream, 1-based orbitals (cip/ai.f:1-42). */
static inline double ai1(const double *mo, int a, int b, int c, int e)
{
    return or_ai(mo, a - 1, b - 1, c - 1, e - 1);
}

/* hntd.f:36-168: <D_a|H|D_b> when the determinants differ by one or two spin-orbitals,
 * 0 when they are identical or differ by more.  Steps: the determinant with more
 * excitations is "1"; its excitation list is reordered to line up with the other's
 * (each swap flips the sign); unmatched holes / particles give the differing
 * spin-orbitals nv1 -> nv2; one difference: F(n1, n2) corrected for the holes and
 * particles of determinant 1; two: (n1 n2|n3 n4) - (n1 n4|n3 n2) by spin. */
double or_hntd(int da, int db, int norb, const int *ne, const int *nd, const int *trou,
               const int *part, const int *iorb, const int *ispin, const double *fock,
               const double *mo)
{
    int ita[20], itb[20], nv1[2] = {0, 0}, nv2[2] = {0, 0}, natu1[2] = {0, 0}, natu2[2] = {0, 0};
    const int kr = (ne[da] - ne[db] <= 0) ? db : da;
    const int kl = (kr == db) ? da : db;
    const int ne1 = ne[kr];
    if (ne1 == 0) return 0.0;
    const int ne2 = ne[kl];
    int nbdif = ne1 - ne2;
    if (nbdif > 2 || ne1 > 10) return 0.0;
    for (int t = 0; t < ne1; ++t) {
        ita[t] = trou[nd[kr] + t];
        ita[t + ne1] = part[nd[kr] + t];
    }
    int ysig = 0;
    if (ne2 > 0) {
        for (int t = 0; t < ne2; ++t) {
            itb[t] = trou[nd[kl] + t];
            itb[t + ne2] = part[nd[kl] + t];
        }
        for (int k = 0; k < 2; ++k) {           /* holes, then particles */
            const int k1 = k * ne1, k2 = k * ne2;
            for (int i = 0; i < ne2; ++i) {
                const int nit1 = itb[k2 + i];
                int found = 0;
                for (int j = k1; j < k1 + ne1; ++j) {
                    if (ita[j] != nit1) continue;
                    if (i != j - k1) {
                        ita[j] = ita[i + k1];
                        ita[i + k1] = nit1;
                        ysig = !ysig;
                    }
                    found = 1;
                    break;
                }
                if (!found) {
                    nbdif = nbdif + 1;
                    if (nbdif > 2) return 0.0;
                    nv1[nbdif - 1] = nit1;
                    natu1[nbdif - 1] = k + 1;
                    natu2[nbdif - 1] = i;
                }
            }
        }
    }
    if (nbdif <= 0) return 0.0;
    const int nar = ne1 - ne2;                  /* extra excitations of determinant 1 */
    for (int i = 0; i < nar; ++i) {
        const int net = ne1 - 1 - i;
        nv1[i] = ita[net];
        nv2[i] = ita[net + ne1];
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
    const int n1 = iorb[nv1[0] - 1], n2 = iorb[nv2[0] - 1];
    const int ns1 = ispin[nv1[0] - 1], ns2 = ispin[nv2[0] - 1];
    double ha = 0.0;
    if (nbdif == 2) {
        const int n3 = iorb[nv1[1] - 1], n4 = iorb[nv2[1] - 1];
        const int ns4 = ispin[nv2[1] - 1];
        if (ns1 == ns2) ha = ai1(mo, n1, n2, n3, n4);
        if (ns1 == ns4) ha = ha - ai1(mo, n1, n4, n3, n2);
    } else {
        ha = fock[tri1s(n1, n2)];
        for (int i = 0; i < ne1; ++i) {
            int np = iorb[ita[i] - 1];
            int nsu = ispin[ita[i] - 1];
            ha = ha - ai1(mo, n1, n2, np, np);
            if (ns1 == nsu) ha = ha + ai1(mo, n1, np, n2, np);
            np = iorb[ita[i + ne1] - 1];
            if (np > norb) continue;
            nsu = ispin[ita[i + ne1] - 1];
            ha = ha + ai1(mo, n1, n2, np, np);
            if (ns1 == nsu) ha = ha - ai1(mo, n1, np, n2, np);
        }
    }
    return ysig ? -ha : ha;
}
