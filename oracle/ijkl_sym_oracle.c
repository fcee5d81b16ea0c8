/*
 * ijkl_sym_oracle.c -- CPU restatement of the reference ijkl for molecules with
 * point-group symmetry (nsym > 1: the non-degenerate groups CS, CI, CN, CNV, CNH, DN, DNH,
 * and the linear groups CINFV, DINFH with their symmetry transformations).
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the GPU path; never linked into the product).
 * Every routine follows the reference loop nests and summation order so that the
 * results match the flang-built reference (oracle/_ref/ijkl_ref) bit for bit:
 *
 *   or_sym_prepare   otodon.f:320-353 (symmetry orbitals of the non-degenerate groups),
 *                    rotat2.f:50-423 (frozen compaction, normalisation of the symmetry
 *                    vectors, classes, degeneracy from tmono.f:37-54, projection,
 *                    normalisation and orthogonality checks, reordering by irrep),
 *                    vector.f:17-46 (per-irrep packing of C)
 *   or_sym_linear    otodon.f:36-144 (linear groups: classes by angular momentum about
 *                    the axis, g/u combinations for DINFH, tables of BLOCKDATA1.f)
 *   or_sym_prepare_so  rotat2 + vector for the table or_sym_linear builds
 *   or_sym_itijkl    itijkl.f:42-298 (class and irrep block loops, the symmetry
 *                    transformations of the linear groups, block directory),
 *                    transd.f:1-379 (label stream -> class-quartet AO block, symmetry
 *                    images), tpqkld.f:1-108 and tijkld.f:1-90 in their general
 *                    (rectangular) forms, with the unit-40 and unit-50 record streams
 *                    kept in memory
 *
 * Conventions: arrays hold the reference's 1-based values; Fortran array a(i,j) of
 * leading dimension d is a[(i-1) + (j-1)*d] here.  No fused multiply-add (built with
 * -ffp-contract=off), so products and sums round as in the reference.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NUMNU(i) (((int64_t)(i) * ((i) - 1)) / 2)

/* ------------------------------------------------------------------ otodon + rotat2 + vector */
/*
 * Inputs (group tables of BLOCKDATA1.f as otodon selects them, otodon.f:281-306):
 *   its[nsym*nsym]      product table its(i,j)
 *   jsym[nsym*nt]       characters jsym(i,l) of irrep i under operation l
 * basis (rondof.f:30-38, 77-82):
 *   ktype[nshell], kloc[nshell] (1-based first AO), newsh[nshell*nt] (1-based),
 *   iptr[3*nt], idtr[6*nt], iftr[10*nt] (diagonal signs, rondof.f:100-118)
 * info records: s_packed, h_packed (N(N+1)/2), c_square (N*N, column-major)
 * namelist: nvecf, ygel[N] (frozen flags), noc, trec, tdeg
 * Outputs:
 *   c_out[N*N]      the MO coefficients rotat2 leaves (written to ijom)
 *   itsyv[N]        irrep of each MO
 *   isymp[nsym], *nsymp   classes (irreps with the same AOs)
 *   ivecsy[nsym*N]  after vector: index of AO i within irrep s's subset, 0 if absent
 *   nbp[nsym], nbo[nsym], ind[nsym+1], c_packed[...]  vector.f's /om/
 * Returns 0, or the error tag rotat2 prints before stopping (2, 1, -1).
 */
static int prepare_impl(int have_so, int norb, int nsym, const int *jsym, int nt, int nshell,
                        const int *ktype, const int *kloc, const int *newsh, const int *iptr,
                        const int *idtr, const int *iftr, const double *s_packed,
                        const double *h_packed, const double *c_square, int nvecf,
                        const unsigned char *ygel, int noc, double trec, double tdeg,
                        double *c_out, int *itsyv, int *isymp, int *nsymp_out, int *ivecsy,
                        int *nbp, int *nbo, int *ind, double *c_packed)
{
    const int n = norb;
    const int64_t nn = (int64_t)n * n;
    int rc = 0;
    /* ---- otodon.f:320-353 (have_so: ivecsy already holds the linear groups' table) */
    if (!have_so) {
        memset(ivecsy, 0, sizeof(int) * (size_t)nsym * n);
        int i1 = 0;
        for (int i = 1; i <= nsym; ++i) {
            const int im1n = (i - 1) * n;
            int i2 = 0;
            for (int j = 1; j <= nshell; ++j) {
                const int kt = ktype[j - 1];
                const int iii = 2 * (kt - 1) + 1;
                for (int k = 1; k <= iii; ++k) {
                    ++i1;
                    if (ivecsy[i1 - 1] != 0) continue;
                    ++i2;
                    for (int l = 1; l <= nt; ++l) {
                        const int k1 = kloc[newsh[(j - 1) + (l - 1) * nshell] - 1] + k - 1 + im1n;
                        int k2 = jsym[(i - 1) + (l - 1) * nsym];
                        if (kt == 2) k2 *= iptr[(k - 1) + 3 * (l - 1)];
                        if (kt == 3) k2 *= idtr[(k - 1) + 6 * (l - 1)];
                        if (kt == 4) k2 *= iftr[(k - 1) + 10 * (l - 1)];
                        const int i3 = i2 * k2;
                        if (ivecsy[k1 - 1] != 0 && ivecsy[k1 - 1] != i3) {
                            for (int ll = 1; ll <= nt; ++ll) {
                                const int k4 = kloc[newsh[(j - 1) + (ll - 1) * nshell] - 1] + k - 1 + im1n;
                                ivecsy[k4 - 1] = 0;
                            }
                            --i2;
                            break;
                        }
                        ivecsy[k1 - 1] = i3;
                    }
                }
            }
        }
    }
    /* ---- rotat2.f:50-60: frozen compaction */
    double *c = (double *)malloc(sizeof(double) * (size_t)nn);
    double *csc = (double *)malloc(sizeof(double) * (size_t)nn);
    double *vecsym = (double *)malloc(sizeof(double) * (size_t)nsym * n);
    double *a = (double *)malloc(sizeof(double) * (size_t)nn);
    double *srec = (double *)malloc(sizeof(double) * (size_t)nsym * n + 8);
    double *ch = (double *)malloc(sizeof(double) * ((size_t)n * (n + 1) / 2 + 1));
    double *cc = (double *)malloc(sizeof(double) * (size_t)nn);
    int *isym = (int *)calloc((size_t)n + 1, sizeof(int));
    int *ideg = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int *iniv = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    int *itr = (int *)malloc(sizeof(int) * ((size_t)nsym + 1));
    memcpy(c, c_square, sizeof(double) * (size_t)nn);
    {
        int nacti = 0;
        for (int i = 1; i <= nvecf; ++i) {
            if (ygel && ygel[i - 1]) continue;
            ++nacti;
            for (int j = 1; j <= n; ++j) c[(int64_t)(nacti - 1) * n + j - 1] = c[(int64_t)(i - 1) * n + j - 1];
        }
    }
    for (int i = 1; i <= nsym; ++i)
        for (int j = 1; j <= n; ++j) vecsym[(i - 1) * n + j - 1] = (double)ivecsy[(i - 1) * n + j - 1];
    for (int i = 1; i <= nvecf; ++i) isym[i] = 0;          /* itsyv of the namelist (0) */
    isymp[0] = 1;
#define SIJ(i, j) s_packed[((i) >= (j) ? NUMNU(i) + (j) : NUMNU(j) + (i)) - 1]
    /* rotat2.f:91-112: normalised symmetry vectors, membership */
    for (int is = 1; is <= nsym; ++is) {
        double rec = 0;
        const int km = (is - 1) * n;
        for (int i = 1; i <= n; ++i)
            for (int j = 1; j <= n; ++j) rec = rec + vecsym[km + i - 1] * SIJ(i, j) * vecsym[km + j - 1];
        if (rec < 1.e-8) rec = 1.0;
        rec = 1.0 / sqrt(rec);
        for (int i = 1; i <= n; ++i) vecsym[km + i - 1] = vecsym[km + i - 1] * rec;
        for (int i = 1; i <= n; ++i) ivecsy[km + i - 1] = fabs(vecsym[km + i - 1]) < 1.e-8 ? 0 : is;
    }
    /* rotat2.f:114-130: classes */
    int nsymp = 1;
    for (int is = 2; is <= nsym; ++is) {
        int found = 0;
        for (int js = 1; js < is && !found; ++js) {
            int same = 1;
            for (int i = 1; i <= n; ++i) {
                const int a_ = ivecsy[(js - 1) * n + i - 1], b_ = ivecsy[(is - 1) * n + i - 1];
                if ((a_ == 0 && b_ != 0) || (b_ == 0 && a_ != 0)) { same = 0; break; }
            }
            if (same) { isymp[is - 1] = isymp[js - 1]; found = 1; }
        }
        if (!found) isymp[is - 1] = ++nsymp;
    }
    /* rotat2.f:131-157: irreps reordered so that each class is contiguous (the product
       table its is left as it was) */
    {
        int iss = 0;
        for (int is = 1; is <= nsymp; ++is)
            for (int js = is; js <= nsym; ++js) {
                if (isymp[js - 1] != is) continue;
                ++iss;
                if (js == iss) continue;
                isymp[js - 1] = isymp[iss - 1];
                isymp[iss - 1] = is;
                for (int i = 1; i <= n; ++i) {
                    if (isym[i] == js) isym[i] = iss;
                    else if (isym[i] == iss) isym[i] = js;
                }
                for (int i = 1; i <= n; ++i) {
                    const int km = (js - 1) * n + i - 1, kj = (iss - 1) * n + i - 1;
                    double t = vecsym[km]; vecsym[km] = vecsym[kj]; vecsym[kj] = t;
                    int u = ivecsy[km]; ivecsy[km] = ivecsy[kj]; ivecsy[kj] = u;
                }
            }
    }
    /* rotat2.f:172-176, tmono.f:37-54: MO diagonal of h for the degeneracy levels */
    for (int i = 1; i <= nvecf; ++i)
        for (int j = 1; j <= n; ++j) {
            double s = 0.0;
            for (int k = 1; k <= n; ++k) {
                const int64_t kj = (k >= j) ? NUMNU(k) + j : NUMNU(j) + k;
                s = s + c[(int64_t)(i - 1) * n + k - 1] * h_packed[kj - 1];
            }
            cc[(int64_t)(i - 1) * n + j - 1] = s;
        }
    {
        int64_t ij = 0;
        for (int i = 1; i <= nvecf; ++i)
            for (int j = 1; j <= i; ++j) {
                double s = 0.0;
                for (int k = 1; k <= n; ++k) s = s + cc[(int64_t)(i - 1) * n + k - 1] * c[(int64_t)(j - 1) * n + k - 1];
                ch[ij++] = s;
            }
    }
    /* rotat2.f:177-203: levels of consecutive MOs with |e_i - e_j| < tdeg */
    int niv = 0;
    {
        int i = 1;
        while (i <= nvecf) {
            int ndeg = 1;
            ++niv;
            int j = i;
            int64_t jj = (int64_t)i * (i + 1) / 2;
            const double cii = ch[jj - 1];
            for (;;) {
                ++j;
                jj += j;
                if (j > nvecf) { ideg[niv] = ndeg; iniv[niv] = i; i = nvecf + 1; break; }
                if (fabs(cii - ch[jj - 1]) < tdeg) { ++ndeg; continue; }
                ideg[niv] = ndeg;
                iniv[niv] = i;
                i += ndeg;
                break;
            }
        }
    }
    /* rotat2.f:216-231: csc = S c */
    for (int i = 1; i <= nvecf; ++i)
        for (int k = 1; k <= n; ++k) {
            double rec = 0.0;
            for (int l = 1; l <= n; ++l) rec = rec + c[(int64_t)(i - 1) * n + l - 1] * SIJ(k, l);
            csc[(int64_t)(i - 1) * n + k - 1] = rec;
        }
    /* rotat2.f:234-323: irrep of every level; projection onto the irreps */
    for (int lv = 1; lv <= niv && !rc; ++lv) {
        const int i = iniv[lv], ndeg = ideg[lv];
        const int64_t mi = (int64_t)(i - 2) * n;
        int ntr = 0;
        for (int is = 1; is <= nsym; ++is) {
            const int mis = (is - 1) * n, mir = (is - 1) * ndeg;
            int yrec = 0;
            for (int j = 1; j <= ndeg; ++j) {
                const int64_t mj = mi + (int64_t)j * n;
                double rec = 0.0;
                for (int l = 1; l <= n; ++l) rec = rec + vecsym[mis + l - 1] * csc[mj + l - 1];
                srec[mir + j - 1] = rec;
                if (fabs(rec) >= trec) yrec = 1;
            }
            if (yrec) itr[++ntr] = is;
        }
        if (ntr > ndeg || (ntr < ndeg && ntr != 1)) {
            rc = ntr < ndeg ? 2 : 1;
        } else if (ntr < ndeg && ntr == 1) {
            for (int t = 1; t <= ndeg; ++t) isym[i + t - 1] = itr[1];
        } else {
            int64_t lk = 0;
            for (int t = 1; t <= ndeg; ++t) {
                const int is = itr[t], mir = (is - 1) * ndeg;
                for (int l = 1; l <= n; ++l) {
                    double aa = 0.0;
                    int64_t jl = mi + l;
                    for (int j = 1; j <= ndeg; ++j) {
                        jl += n;
                        aa = aa + srec[mir + j - 1] * c[jl - 1];
                    }
                    a[lk++] = aa;
                }
                isym[i + t - 1] = is;
            }
            for (int64_t k = 0; k < (int64_t)ndeg * n; ++k) c[mi + n + k] = a[k];
        }
    }
    /* rotat2.f:327-383: normalisation and orthogonality */
    if (!rc) {
        for (int i = 1; i <= nvecf; ++i)
            for (int k = 1; k <= n; ++k) {
                double rec = 0.0;
                for (int l = 1; l <= n; ++l) rec = rec + c[(int64_t)(i - 1) * n + l - 1] * SIJ(k, l);
                csc[(int64_t)(i - 1) * n + k - 1] = rec;
            }
        for (int i = 1; i <= nvecf && !rc; ++i) {
            const int64_t mi = (int64_t)(i - 1) * n;
            double rec = 0.0;
            for (int k = 1; k <= n; ++k) rec = rec + c[mi + k - 1] * csc[mi + k - 1];
            double rec0 = sqrt(fabs(rec));
            if (rec0 < 1.e-6) { rc = -1; break; }
            rec0 = 1.0 / rec0;
            for (int k = 1; k <= n; ++k) {
                csc[mi + k - 1] = csc[mi + k - 1] * rec0;
                c[mi + k - 1] = rec0 * c[mi + k - 1];
            }
        }
        for (int i = 2; i <= nvecf && !rc; ++i)
            for (int j = 1; j < i; ++j) {
                double rec = 0.0;
                for (int k = 1; k <= n; ++k)
                    rec = rec + c[(int64_t)(j - 1) * n + k - 1] * csc[(int64_t)(i - 1) * n + k - 1];
                if (fabs(rec) > trec) { rc = 1; break; }
            }
    }
    if (!rc) {
        /* rotat2.f:386-420: occupied MOs by irrep, then virtual MOs by irrep */
        int koc = 0, kvirt = noc;
        for (int is = 1; is <= nsym; ++is)
            for (int i = 1; i <= nvecf; ++i) {
                if (isym[i] != is) continue;
                const int k = i > noc ? ++kvirt : ++koc;
                memcpy(csc + (int64_t)(k - 1) * n, c + (int64_t)(i - 1) * n, sizeof(double) * n);
                itsyv[k - 1] = isym[i];
            }
        memcpy(c, csc, sizeof(double) * (size_t)nvecf * n);
        memcpy(c_out, c, sizeof(double) * (size_t)nn);
        /* vector.f:17-46 */
        int lve = 0;
        for (int ls = 1; ls <= nsym; ++ls) {
            ind[ls - 1] = lve + 1;
            int k = 0;
            for (int i = 1; i <= n; ++i)
                if (ivecsy[(ls - 1) * n + i - 1] != 0) ivecsy[(ls - 1) * n + i - 1] = ++k;
            nbp[ls - 1] = k;
            k = 0;
            for (int i = 1; i <= nvecf; ++i) {
                if (itsyv[i - 1] != ls) continue;
                ++k;
                for (int j = 1; j <= n; ++j)
                    if (ivecsy[(ls - 1) * n + j - 1] != 0) c_packed[lve++] = c[(int64_t)(i - 1) * n + j - 1];
            }
            nbo[ls - 1] = k;
        }
        ind[nsym] = lve + 1;
    }
#undef SIJ
    *nsymp_out = nsymp;
    free(c); free(csc); free(vecsym); free(a); free(srec); free(ch); free(cc);
    free(isym); free(ideg); free(iniv); free(itr);
    return rc;
}

int or_sym_prepare(int norb, int nsym, const int *its, const int *jsym, int nt, int nshell,
                   const int *ktype, const int *kloc, const int *newsh, const int *iptr,
                   const int *idtr, const int *iftr, const double *s_packed,
                   const double *h_packed, const double *c_square, int nvecf,
                   const unsigned char *ygel, int noc, double trec, double tdeg,
                   double *c_out, int *itsyv, int *isymp, int *nsymp_out, int *ivecsy,
                   int *nbp, int *nbo, int *ind, double *c_packed)
{
    (void)its;
    return prepare_impl(0, norb, nsym, jsym, nt, nshell, ktype, kloc, newsh, iptr, idtr, iftr,
                        s_packed, h_packed, c_square, nvecf, ygel, noc, trec, tdeg, c_out, itsyv,
                        isymp, nsymp_out, ivecsy, nbp, nbo, ind, c_packed);
}

/* rotat2 + vector for a symmetry-orbital table so[nsym*N] that otodon built already (the
   linear groups, or_sym_linear). */
int or_sym_prepare_so(int norb, int nsym, const int *so, const double *s_packed,
                      const double *h_packed, const double *c_square, int nvecf,
                      const unsigned char *ygel, int noc, double trec, double tdeg,
                      double *c_out, int *itsyv, int *isymp, int *nsymp_out, int *ivecsy,
                      int *nbp, int *nbo, int *ind, double *c_packed)
{
    memcpy(ivecsy, so, sizeof(int) * (size_t)nsym * norb);
    return prepare_impl(1, norb, nsym, NULL, 0, 0, NULL, NULL, NULL, NULL, NULL, NULL, s_packed,
                        h_packed, c_square, nvecf, ygel, noc, trec, tdeg, c_out, itsyv, isymp,
                        nsymp_out, ivecsy, nbp, nbo, ind, c_packed);
}

/* ------------------------------------------------------------------ linear groups */
/* BLOCKDATA1.f:15-41: product tables itsdh (14 x 14) and itscv (7 x 7) and the symmetry
   transformations isyded (14 x 2) and isydec (7 x 2), column-major as the DATA statements
   list them. */
static const int kItsdh[14 * 14] = {
    1, 2, 3, 4, 5, 6, 1, 2, 9, 10, 3, 4, 5, 6,    2, 1, 4, 3, 6, 5, 2, 1, 10, 9, 4, 3, 6, 5,
    3, 4, 1, 2, 9, 10, 3, 4, 5, 6, 1, 2, 9, 10,   4, 3, 2, 1, 10, 9, 4, 3, 6, 5, 2, 1, 10, 9,
    5, 6, 9, 10, 1, 2, 5, 6, 3, 4, 9, 10, 1, 2,   6, 5, 10, 9, 2, 1, 6, 5, 4, 3, 10, 9, 2, 1,
    1, 2, 3, 4, 5, 6, 1, 2, 9, 10, 3, 4, 5, 6,    2, 1, 4, 3, 6, 5, 2, 1, 10, 9, 4, 3, 6, 5,
    9, 10, 5, 6, 3, 4, 9, 10, 1, 2, 5, 6, 3, 4,   10, 9, 6, 5, 4, 3, 10, 9, 2, 1, 6, 5, 4, 3,
    3, 4, 1, 2, 9, 10, 3, 4, 5, 6, 1, 2, 9, 10,   4, 3, 2, 1, 10, 9, 4, 3, 6, 5, 2, 1, 10, 9,
    5, 6, 9, 10, 1, 2, 5, 6, 3, 4, 9, 10, 1, 2,   6, 5, 10, 9, 2, 1, 6, 5, 4, 3, 10, 9, 2, 1};
static const int kItscv[7 * 7] = {
    1, 2, 3, 1, 5, 2, 3,   2, 1, 5, 2, 3, 1, 5,   3, 5, 1, 3, 2, 5, 1,   1, 2, 3, 1, 5, 2, 3,
    5, 3, 2, 5, 1, 3, 2,   2, 1, 5, 2, 3, 1, 5,   3, 5, 1, 3, 2, 5, 1};
static const int kIsyded[14 * 2] = {1, 2, 5, 6, -3, -4, -7, -8, -9, -10, 13, 14, -11, -12,
                                    1, 2, 0, 0, 0, 0, 9, 10, -7, -8, 0, 0, 0, 0};
static const int kIsydec[7 * 2] = {1, 3, -2, -4, -5, 7, -6,   1, 0, 0, 5, -4, 0, 0};
/* otodon.f:37-39: component of a shell of type ktype carrying class jnk (0 = sigma,
   1 = pi_x, 2 = pi_y, 3 = delta_x, 4 = delta_y, 5, 6 = phi), -1 if none:
   idinf2(ktype + 4 jnk). */
static const int kIdinf2[28] = {0, 2, 0, 0, -1, 0, 1, 1, -1, 1, 2, 2, -1, -1, 3, 3,
                                -1, -1, 4, 4, -1, -1, -1, 5, -1, -1, -1, 6};

/*
 * otodon.f:36-144 for CINFV (dinfh = 0) and DINFH (dinfh = 1): nsym (5 or 7, 10 or 14 with
 * f shells), the product table its (nsym x nsym, column-major), the number of symmetry
 * transformations ntrans (one each for the p and the d shells present, starting from the
 * 0 that ijkl.f:48 passes through rondof) with isydeg (nsym x ntrans, column-major), and
 * the signed symmetry-orbital table so[nsym*N] (otodon's ivecsy).  Per shell: ktype,
 * kloc (1-based first AO), zsh (the z coordinate of the shell's atom) and newsh2 (1-based
 * image shell under the second operation of the group pshf ran in, D-inf-h only).
 * The branch otodon.f:104-106 / 139-142 (no p shell and ntrans = 1) cannot be taken
 * with ntrans entering as 0, so it has no counterpart here.
 */
int or_sym_linear(int dinfh, int norb, int nshell, const int *ktype, const int *kloc,
                  const double *zsh, const int *newsh2, int *nsym_out, int *its,
                  int *ntrans_out, int *isydeg, int *so)
{
    int yp = 0, yd = 0, yf = 0;
    for (int i = 0; i < nshell; ++i) {
        if (ktype[i] == 2) yp = 1;
        if (ktype[i] == 3) yd = 1;
        if (ktype[i] == 4) yf = 1;
    }
    if (yf) { yp = 1; yd = 1; }
    if (yd) yp = 1;
    const int ntrans = (yp ? 1 : 0) + (yd ? 1 : 0);
    const int nsym = dinfh ? (yf ? 14 : 10) : (yf ? 7 : 5);
    memset(so, 0, sizeof(int) * (size_t)nsym * norb);
    if (dinfh) {
        int yg = 0, ish = -norb;
        for (int ink = 1; ink <= nsym; ++ink) {
            yg = !yg;
            ish += norb;
            const int jnk = (ink - 1) / 2;
            int i2 = 0;
            for (int i = 1; i <= nshell; ++i) {
                const int idp = kIdinf2[ktype[i - 1] + jnk * 4 - 1];
                if (idp < 0) continue;
                int i1 = ish + kloc[i - 1] + idp;
                const int yki = (ktype[i - 1] - 2 * (ktype[i - 1] / 2)) == 0;
                const int ysym = yki != yg;
                if (!ysym && fabs(zsh[i - 1]) < 1.e-8) continue;
                if (so[i1 - 1] != 0) continue;
                ++i2;
                so[i1 - 1] = i2;
                if (fabs(zsh[i - 1]) > 1.e-8) {
                    i1 = ish + kloc[newsh2[i - 1] - 1] + idp;
                    if (so[i1 - 1] == 0) so[i1 - 1] = ysym ? i2 : -i2;
                }
            }
        }
    } else {
        int ish = -norb;
        for (int ink = 1; ink <= nsym; ++ink) {
            const int jnk = ink - 1;
            int i2 = 0;
            ish += norb;
            for (int i = 1; i <= nshell; ++i) {
                const int idp = kIdinf2[ktype[i - 1] + jnk * 4 - 1];
                if (idp < 0) continue;
                const int i1 = ish + kloc[i - 1] + idp;
                if (so[i1 - 1] != 0) continue;
                ++i2;
                so[i1 - 1] = i2;
            }
        }
    }
    const int *tab = dinfh ? kItsdh : kItscv;
    const int *deg = dinfh ? kIsyded : kIsydec;
    const int ld = dinfh ? 14 : 7;
    for (int j = 0; j < nsym; ++j) {
        for (int i = 0; i < nsym; ++i) its[i + j * nsym] = tab[i + j * ld];
    }
    for (int t = 0; t < ntrans; ++t)
        for (int j = 0; j < nsym; ++j) isydeg[j + t * nsym] = deg[j + t * ld];
    *nsym_out = nsym;
    *ntrans_out = ntrans;
    return 0;
}

/* ------------------------------------------------------------------ itijkl and its routines */
typedef struct {
    int norb, nsym, nt;
    const int *its, *isymp, *nbp, *nbo, *ind, *ivecsy;
    const double *c;                   /* /om/ c, packed by vector.f */
    const int *ishell, *kloc, *ktyp, *numc, *newsh, *iptr, *idtr, *iftr;
    int nshell;
    const double *xx;
    const int32_t *ix;
    int64_t nlab;
    /* unit-40 and unit-50 streams */
    double *h;
    int64_t nh, caph;
    double *out;
    int64_t nout, capout;
    int oom;
} SymState;

static void put_h(SymState *S, double v) {
    if (S->nh == S->caph) {
        S->caph = S->caph ? 2 * S->caph : 1 << 16;
        double *p = (double *)realloc(S->h, sizeof(double) * (size_t)S->caph);
        if (!p) { S->oom = 1; return; }
        S->h = p;
    }
    S->h[S->nh++] = v;
}

static void put_out(SymState *S, double v) {
    if (S->nout < S->capout) S->out[S->nout] = v;
    else S->oom = 2;
    S->nout++;
}

/* transd.f:1-379.  Returns the last store index (0 = nothing stored, transd.f:376). */
static int64_t sym_transd(SymState *S, int msi, int msj, int msk, int msl, double *b) {
    const int n = S->norb, nt = S->nt;
    const int nbpj = S->nbp[msj - 1], nbpl = S->nbp[msl - 1];
    const int64_t nbpkl = (int64_t)S->nbp[msk - 1] * S->nbp[msl - 1];
    const int64_t nbpkls = (nbpl + nbpkl) / 2;
    int *nsp = (int *)calloc(4 * ((size_t)n + 1), sizeof(int));
    unsigned char *yfl = (unsigned char *)calloc(8 * ((size_t)n + 1), 1);
    int *nspi = nsp, *nspj = nsp + n + 1, *nspk = nsp + 2 * (n + 1), *nspl = nsp + 3 * (n + 1);
    unsigned char *yis = yfl, *yjs = yfl + (n + 1), *yks = yfl + 2 * (n + 1), *yls = yfl + 3 * (n + 1);
    unsigned char *yij = yfl + 4 * (n + 1), *ykl = yfl + 5 * (n + 1), *yijkls = yfl + 6 * (n + 1);
    int ki = 0, kj = 0, kk = 0, kl = 0;
    for (int i = 1; i <= n; ++i) {
        yis[i] = S->ivecsy[(msi - 1) * n + i - 1] == 0;
        yjs[i] = S->ivecsy[(msj - 1) * n + i - 1] == 0;
        yks[i] = S->ivecsy[(msk - 1) * n + i - 1] == 0;
        yls[i] = S->ivecsy[(msl - 1) * n + i - 1] == 0;
        yij[i] = yis[i] && yjs[i];
        ykl[i] = yks[i] && yls[i];
        if (!yis[i]) nspi[i] = ++ki;
        if (!yjs[i]) nspj[i] = ++kj;
        if (!yks[i]) nspk[i] = ++kk;
        if (!yls[i]) nspl[i] = ++kl;
        yijkls[i] = yij[i] && ykl[i];
    }
    const int *isymp = S->isymp;
    const int ypq = isymp[msi - 1] == isymp[msj - 1];
    const int yrs = isymp[msk - 1] == isymp[msl - 1];
    const int ypqrs = isymp[msi - 1] == isymp[msk - 1] && isymp[msj - 1] == isymp[msl - 1];
    const int ypqrst = ypq && ypqrs;
    const int ijcas = 1 + (ypq ? 1 : 0) + (yrs ? 2 : 0);
    const double dfnt = 1.0 / (double)nt;
    int64_t ijkl = 0;
    double qq4 = 1.0;
    int icd = 0, jcd = 0, kcd = 0, lcd = 0;
    int ysy[64];
    for (int x = 0; x < 64; ++x) ysy[x] = 1;
#define IA(x) NUMNU(x)
#define STORE(pos) do { ijkl = (pos); b[ijkl - 1] = val; } while (0)
    for (int64_t m = 0; m < S->nlab; ++m) {
        const uint32_t lab = (uint32_t)S->ix[m];
        if (lab == 0) break;
        int i = (int)(lab >> 24), j = (int)((lab >> 16) & 255u), k = (int)((lab >> 8) & 255u),
            l = (int)(lab & 255u);
        if (yijkls[i] || yijkls[j] || yijkls[k] || yijkls[l]) continue;
        int icas;
        if (ypqrst) icas = 8;
        else if (yij[i] || yij[j]) {
            if (ykl[i] || ykl[j]) continue;
            if (yij[k] || yij[l]) continue;
            icas = 5;
        } else if (ykl[k] || ykl[l]) {
            if (yij[k] || yij[l]) continue;
            if (ykl[i] || ykl[j]) continue;
            icas = 5;
        } else if (ypqrs) icas = 7;
        else if (yij[k] || yij[l] || ykl[i] || ykl[j]) icas = ijcas;
        else icas = 6;
        /* transd.f:125-209: symmetry images of the shell quartet */
        const int ig = S->ishell[i - 1], jg = S->ishell[j - 1], kg = S->ishell[k - 1], lg = S->ishell[l - 1];
        int nct[4] = {1, 1, 1, 1}, ncn[4] = {0, 0, 0, 0};
        if (nt >= 2) {
            int ic = ig, jc = jg, kc = kg, lc = lg;
            if (!(ic > jc)) { jc = ig; ic = jg; }
            if (!(kc > lc)) { kc = lg; lc = kg; }
            if (ic < kc || (ic == kc && lc > jc)) {
                int t = ic; ic = kc; kc = t;
                t = jc; jc = lc; lc = t;
            }
            nct[0] = S->ktyp[i - 1]; nct[1] = S->ktyp[j - 1]; nct[2] = S->ktyp[k - 1]; nct[3] = S->ktyp[l - 1];
            ncn[0] = S->numc[i - 1]; ncn[1] = S->numc[j - 1]; ncn[2] = S->numc[k - 1]; ncn[3] = S->numc[l - 1];
            if (lc != lcd || kc != kcd || jc != jcd || ic != icd) {
                int nq4 = 1, ngen = 0;
                int mti[64], mtj[64], mtk[64], mtl[64];
                for (int g = 2; g <= nt; ++g) {
                    ysy[g] = 1;
                    const int it = S->newsh[(ic - 1) + (g - 1) * S->nshell];
                    const int jt = S->newsh[(jc - 1) + (g - 1) * S->nshell];
                    const int kt = S->newsh[(kc - 1) + (g - 1) * S->nshell];
                    const int lt = S->newsh[(lc - 1) + (g - 1) * S->nshell];
                    int n1, n2, n3, n4;
                    if (it < jt) { n1 = jt; n2 = it; } else { n1 = it; n2 = jt; }
                    if (kt < lt) { n3 = lt; n4 = kt; } else { n3 = kt; n4 = lt; }
                    if (n1 < n3 || (n1 == n3 && n2 < n4)) {
                        int t = n1; n1 = n3; n3 = t;
                        t = n2; n2 = n4; n4 = t;
                    }
                    if (n1 == ic && n2 == jc && n3 == kc && n4 == lc) { ++nq4; ysy[g] = 0; continue; }
                    int dup = 0;
                    for (int x = 0; x < ngen; ++x)
                        if (n1 == mti[x] && n2 == mtj[x] && n3 == mtk[x] && n4 == mtl[x]) { dup = 1; break; }
                    if (dup) { ysy[g] = 0; continue; }
                    mti[ngen] = n1; mtj[ngen] = n2; mtk[ngen] = n3; mtl[ngen] = n4;
                    ++ngen;
                }
                lcd = lc; kcd = kc; jcd = jc; icd = ic;
                qq4 = (double)nq4 * dfnt;
            }
        }
        /* transd.f:210-218 */
        double val = S->xx[m] * qq4;
        const int yijgr = i == j, yklgr = k == l;
        if (yijgr) val = val + val;
        if (yklgr) val = val + val;
        const int yijklg = i == k && j == l;
        if (yijklg) val = val + val;
        const double vali = val;
        for (int indx = 1; indx <= nt; ++indx) {
            int yijg = yijgr, yklg = yklgr;
            if (indx > 1) {
                if (!ysy[indx]) continue;
                int ysig = 0;
                for (int mn = 0; mn < 4; ++mn) {
                    if (nct[mn] == 1) continue;
                    const int mcn = ncn[mn] + 1;
                    int sgn;
                    if (nct[mn] < 3) sgn = S->iptr[(mcn - 1) + 3 * (indx - 1)];
                    else if (nct[mn] == 3) sgn = S->idtr[(mcn - 1) + 6 * (indx - 1)];
                    else sgn = S->iftr[(mcn - 1) + 10 * (indx - 1)];
                    if (sgn < 0) ysig = !ysig;
                }
                val = ysig ? -vali : vali;
                i = S->kloc[S->newsh[(ig - 1) + (indx - 1) * S->nshell] - 1] + ncn[0];
                j = S->kloc[S->newsh[(jg - 1) + (indx - 1) * S->nshell] - 1] + ncn[1];
                k = S->kloc[S->newsh[(kg - 1) + (indx - 1) * S->nshell] - 1] + ncn[2];
                l = S->kloc[S->newsh[(lg - 1) + (indx - 1) * S->nshell] - 1] + ncn[3];
            } else {
                val = vali;
            }
            /* transd.f:245-373: the stores of the case; yflem asks for a second pass with
               the pairs exchanged (940) */
            int yflem = 0;
            int cas = icas;
            if (cas == 5) {                         /* 940 */
                int t = i; i = k; k = t;
                t = j; j = l; l = t;
                t = yijg; yijg = yklg; yklg = t;
                cas = ijcas;
            } else if (cas == 6) {                  /* 980 */
                yflem = !yijklg;
                cas = ijcas;
            }
            for (;;) {
                if (cas == 1) {                     /* 900 */
                    int64_t ij1 = 0, ij2 = 0, kl1;
                    int ok = 1;
                    if (yis[i] || yjs[j]) {
                        if (yis[j] || yjs[i]) ok = 0;
                        else ij1 = (int64_t)nbpj * (nspi[j] - 1) + nspj[i];
                    } else {
                        ij1 = (int64_t)nbpj * (nspi[i] - 1) + nspj[j];
                        if (!(yis[j] || yjs[i] || yijg)) ij2 = (int64_t)nbpj * (nspi[j] - 1) + nspj[i];
                    }
                    if (ok) {
                        if (!(yks[k] || yls[l])) {
                            kl1 = (int64_t)nbpl * (nspk[k] - 1) + nspl[l];
                            STORE(nbpkl * (ij1 - 1) + kl1);
                            if (ij2) STORE(nbpkl * (ij2 - 1) + kl1);
                        }
                        if (!(yks[l] || yls[k] || yklg)) {
                            kl1 = (int64_t)nbpl * (nspk[l] - 1) + nspl[k];
                            STORE(nbpkl * (ij1 - 1) + kl1);
                            if (ij2) STORE(nbpkl * (ij2 - 1) + kl1);
                        }
                    }
                } else if (cas == 2) {              /* 910 */
                    const int64_t ij1 = i >= j ? IA(nspi[i]) + nspj[j] : IA(nspj[j]) + nspi[i];
                    int skip914 = 0;
                    if (!(yks[k] || yls[l])) {
                        const int64_t kl1 = (int64_t)nbpl * (nspk[k] - 1) + nspl[l];
                        STORE(nbpkl * (ij1 - 1) + kl1);
                        if (yklg) skip914 = 1;
                    }
                    if (!skip914 && !(yls[k] || yks[l])) {
                        const int64_t kl1 = (int64_t)nbpl * (nspk[l] - 1) + nspl[k];
                        STORE(nbpkl * (ij1 - 1) + kl1);
                    }
                } else if (cas == 3) {              /* 920 */
                    const int64_t kl1 = k >= l ? IA(nspk[k]) + nspl[l] : IA(nspl[l]) + nspk[k];
                    int skip924 = 0;
                    if (!(yis[i] || yjs[j])) {
                        const int64_t ij1 = (int64_t)nbpj * (nspi[i] - 1) + nspj[j];
                        STORE(nbpkls * (ij1 - 1) + kl1);
                        if (yijg) skip924 = 1;
                    }
                    if (!skip924 && !(yis[j] || yjs[i])) {
                        const int64_t ij1 = (int64_t)nbpj * (nspi[j] - 1) + nspj[i];
                        STORE(nbpkls * (ij1 - 1) + kl1);
                    }
                } else if (cas == 4) {              /* 930 */
                    const int64_t ij1 = i >= j ? IA(nspi[i]) + nspj[j] : IA(nspj[j]) + nspi[i];
                    const int64_t kl1 = k >= l ? IA(nspk[k]) + nspl[l] : IA(nspl[l]) + nspk[k];
                    STORE(nbpkls * (ij1 - 1) + kl1);
                } else if (cas == 7) {              /* 990 */
                    yflem = 1;
                    for (;;) {                      /* 991 .. 999 */
                        int to235 = 0;
                        if (!(yis[i] || yjs[j])) {
                            const int64_t ij1 = (int64_t)nbpj * (nspi[i] - 1) + nspj[j];
                            int to999 = 0;
                            if (!(yks[k] || yls[l])) {
                                const int64_t kl1 = (int64_t)nbpl * (nspk[k] - 1) + nspl[l];
                                if (k < i || (k == i && j >= l)) STORE(ij1 * (ij1 - 1) / 2 + kl1);
                                else STORE(kl1 * (kl1 - 1) / 2 + ij1);
                                if (yklg) to999 = 1;
                            }
                            if (!to999 && !(yks[l] || yls[k])) {
                                if (!yflem && yijklg) { to235 = 1; }
                                else {
                                    const int64_t kl1 = (int64_t)nbpl * (nspk[l] - 1) + nspl[k];
                                    if (l < i || (l == i && j >= k)) STORE(ij1 * (ij1 - 1) / 2 + kl1);
                                    else STORE(kl1 * (kl1 - 1) / 2 + ij1);
                                }
                            }
                        }
                        if (to235) break;
                        /* 999 */
                        if (!yflem || yijg) break;
                        yflem = 0;
                        if (yis[j] || yjs[i]) break;
                        int t = i; i = j; j = t;
                        t = k; k = l; l = t;
                    }
                    yflem = 0;
                } else {                            /* 1000: icas 8 */
                    if (indx != 1) {
                        if (i < j) { int t = i; i = j; j = t; }
                        if (k < l) { int t = k; k = l; l = t; }
                        if (i < k || (i == k && j < l)) {
                            int t = i; i = k; k = t;
                            t = j; j = l; l = t;
                        }
                    }
                    const int64_t ij1 = IA(nspi[i]) + nspj[j];
                    const int64_t kl1 = IA(nspk[k]) + nspl[l];
                    STORE(ij1 * (ij1 - 1) / 2 + kl1);
                }
                /* 1050 */
                if (!yflem) break;
                {                                   /* 940 */
                    int t = i; i = k; k = t;
                    t = j; j = l; l = t;
                    t = yijg; yijg = yklg; yklg = t;
                    yflem = 0;
                    cas = ijcas;
                }
            }
        }
    }
#undef STORE
#undef IA
    free(nsp);
    free(yfl);
    return ijkl;
}

/* tpqkld.f:1-108 (general form), appending to the unit-40 stream. */
static void sym_tpqkld(SymState *S, int mp, int mq, int mr, int ms, int mk, int ml, int ypq,
                       int yrs, int ypqrs, int ykl, int kc, int lc, const double *b, double *tab,
                       double *tac) {
    (void)mk;
    const double *c = S->c;
    const int np = mp, nr = mr, nl = ml;
    int nq = mq, ns = ms, nk = mk;
    const int64_t mpqm = ypq ? NUMNU(np + 1) : (int64_t)np * mq;
    int64_t npq = 0;
    for (int p = 1; p <= np; ++p) {
        int64_t mpq = 0;
        if (ypq) nq = p;
        for (int q = 1; q <= nq; ++q) {
            if (!ypq) { npq = mpq + p; mpq += np; }
            else ++npq;
            int64_t nrs = 0, mij = 0, sr = 0;
            for (int r = 1; r <= nr; ++r) {
                int64_t mrs = 0, rs = 0;
                if (yrs) ns = r;
                for (int s = 1; s <= ns; ++s) {
                    if (!yrs) {
                        nrs = mrs + r;
                        mrs += nr;
                        if (!ypqrs) {
                            mij = mpqm * (nrs - 1);
                            tab[sr + s - 1] = b[mij + npq - 1];
                        } else if (npq >= nrs) tab[sr + s - 1] = b[NUMNU(npq) + nrs - 1];
                        else tab[sr + s - 1] = b[NUMNU(nrs) + npq - 1];
                    } else {
                        ++nrs;
                        double pqrs;
                        if (!ypqrs) { pqrs = b[mij + npq - 1]; mij += mpqm; }
                        else if (npq >= nrs) pqrs = b[NUMNU(npq) + nrs - 1];
                        else pqrs = b[NUMNU(nrs) + npq - 1];
                        tab[rs + r - 1] = pqrs;
                        tab[sr + s - 1] = pqrs;
                    }
                    rs += nr;
                }
                sr += ms;
            }
            int64_t sl = lc - 1;
            for (int l = 1; l <= nl; ++l) {
                int64_t srr = 0;
                for (int r = 1; r <= nr; ++r) {
                    double sum = 0.0;
                    for (int s = 1; s <= ms; ++s) sum = sum + c[sl + s - 1] * tab[srr + s - 1];
                    srr += ms;
                    tac[r - 1] = sum;
                }
                if (ykl) nk = l;
                int64_t rk = kc - 1;
                for (int k = 1; k <= nk; ++k) {
                    double sum = 0.0;
                    for (int r = 1; r <= nr; ++r) sum = sum + c[rk + r - 1] * tac[r - 1];
                    rk += nr;
                    put_h(S, sum);
                }
                sl += ms;
            }
        }
    }
}

/* tijkld.f:1-90 (general form), appending to the unit-50 stream. */
static void sym_tijkld(SymState *S, int mi, int mj, int mk, int ml, int ic, int jc, int yij,
                       int ykl, int yijkl, int mp, int mq, int ypq, const double *b, double *tab,
                       double *tac) {
    const double *c = S->c;
    int nk = mk, ni = mi, nq = mq;
    const int64_t mkl = ykl ? NUMNU(mk + 1) : (int64_t)mk * ml;
    int64_t ilk = 0;
    for (int l = 1; l <= ml; ++l) {
        if (ykl) nk = l;
        for (int k = 1; k <= nk; ++k) {
            ++ilk;
            int64_t nkl = 0, qp = 0;
            for (int p = 1; p <= mp; ++p) {
                if (ypq) nq = p;
                for (int q = 1; q <= nq; ++q) {
                    tab[qp + q - 1] = b[nkl + ilk - 1];
                    nkl += mkl;
                }
                if (ypq) {
                    int64_t pq = 0;
                    for (int q = 1; q <= nq; ++q) {
                        tab[pq + p - 1] = tab[qp + q - 1];
                        pq += mp;
                    }
                }
                qp += mq;
            }
            int64_t nij = 0, qj = jc - 1;
            int done = 0;
            for (int j = 1; j <= mj && !done; ++j) {
                int64_t qpp = 0;
                for (int p = 1; p <= mp; ++p) {
                    double sum = 0.0;
                    for (int q = 1; q <= mq; ++q) sum = sum + c[qj + q - 1] * tab[qpp + q - 1];
                    qpp += mq;
                    tac[p - 1] = sum;
                }
                if (yij) ni = j;
                int64_t pi = ic - 1;
                for (int i = 1; i <= ni; ++i) {
                    ++nij;
                    double sum = 0.0;
                    for (int p = 1; p <= mp; ++p) sum = sum + c[pi + p - 1] * tac[p - 1];
                    put_out(S, sum);
                    if (yijkl && nij == ilk) { done = 1; break; }
                    pi += mp;
                }
                qj += mq;
            }
        }
    }
}

/* itijkl.f:28-36: ltabs(nklst, ktypsy), the type of a block that a symmetry
   transformation maps onto an earlier block (column-major as the DATA statement). */
static const int kLtabs[100] = {
    1, 1, 1, 1, 1, 1, 1, 1, 1, 10,   1, 2, 3, 4, 5, 6, 7, 8, 9, 1,
    1, 3, 2, 5, 4, 7, 6, 9, 8, 1,    1, 4, 5, 2, 3, 8, 9, 6, 7, 1,
    1, 5, 4, 3, 2, 9, 8, 7, 6, 1,    1, 6, 7, 8, 9, 2, 3, 4, 5, 1,
    1, 7, 6, 9, 8, 3, 2, 5, 4, 1,    1, 8, 9, 6, 7, 4, 5, 2, 3, 1,
    1, 9, 8, 7, 6, 5, 4, 3, 2, 1,    10, 1, 1, 1, 1, 1, 1, 1, 1, 1};

/*
 * itijkl.f:42-298 for nsym > 1.  With ntrsy > 0 (the linear groups), the symmetry
 * transformations isydeg (nsym x ntrsy, column-major) map irrep quartets onto earlier
 * ones (itijkl.f:121-200): such a block is not computed, its directory entry points at
 * the earlier block with a sign and a permutation type, and a block the transformation
 * maps onto itself with the pairs exchanged is stored as a triangle.
 * Inputs: the /om/ and /hondo/ state after or_sym_prepare and rondof (ishell, kloc,
 * ktyp, numc per AO, 1-based; newsh per shell and operation), the label stream.
 * Outputs: out[] the f50 MO stream in tijkld's order (up to out_cap values; *n_out gets
 * the count), ipair[] (2 per block, up to pair_cap blocks; *npaire), nad[nsyt] and
 * ktypsy[nsyt] (nsyt = min(2000, P(P(nsym)))).  Returns 0, or -1 on allocation failure /
 * too small an output.
 */
int or_sym_itijkl(int norb, int nsym, const int *its, const int *isymp, const int *nbp,
                  const int *nbo, const int *ind, const int *ivecsy, const double *c_packed,
                  int nt, int nshell, const int *ishell, const int *kloc, const int *ktyp,
                  const int *numc, const int *newsh, const int *iptr, const int *idtr,
                  const int *iftr, int ntrsy, const int *isydeg, const double *xx,
                  const int32_t *ix, int64_t nlab, double *out, int64_t out_cap, int64_t *n_out,
                  int *ipair, int pair_cap, int *npaire, int *nad, int *ktypsy)
{
    SymState S;
    memset(&S, 0, sizeof(S));
    S.norb = norb; S.nsym = nsym; S.nt = nt; S.its = its; S.isymp = isymp; S.nbp = nbp;
    S.nbo = nbo; S.ind = ind; S.ivecsy = ivecsy; S.c = c_packed; S.ishell = ishell;
    S.kloc = kloc; S.ktyp = ktyp; S.numc = numc; S.newsh = newsh; S.iptr = iptr;
    S.idtr = idtr; S.iftr = iftr; S.nshell = nshell; S.xx = xx; S.ix = ix; S.nlab = nlab;
    S.out = out; S.capout = out_cap;
    int nsyt = nsym * (nsym + 1) / 2;
    nsyt = nsyt * (nsyt + 1) / 2;
    if (nsyt > 2000) nsyt = 2000;
    for (int i = 0; i < nsyt; ++i) { ktypsy[i] = 1; nad[i] = 0; }
    int npair = 0;
    const int nsymp = isymp[nsym - 1];
    int64_t maxblk = 1;                       /* largest class-quartet AO block */
    for (int a = 0; a < nsym; ++a)
        for (int b2 = 0; b2 < nsym; ++b2) {
            const int64_t x = (int64_t)nbp[a] * nbp[b2];
            if (x * x > maxblk) maxblk = x * x;
        }
    double *blk = (double *)malloc(sizeof(double) * (size_t)maxblk);
    double *tab = (double *)malloc(sizeof(double) * ((size_t)norb * norb + 1));
    double *tac = (double *)malloc(sizeof(double) * ((size_t)norb + 1));
    struct Rec { int64_t npqkl; int mi, mj, mk, ml, ic, jc; int yij, ykl, yijkl; } *rec =
        (struct Rec *)malloc(sizeof(struct Rec) * 4096);
    int rc = 0;
    if (!blk || !tab || !tac || !rec) rc = -1;
#define ITS(a, b) its[((a) - 1) + ((b) - 1) * nsym]
#define ISYDEG(a, t) isydeg[((a) - 1) + ((t) - 1) * nsym]
    for (int msi = 1; msi <= nsymp && !rc; ++msi)
        for (int msj = 1; msj <= nsymp && !rc; ++msj) {
            const int ypq = msi == msj;
            for (int msk = 1; msk <= nsymp && !rc; ++msk)
                for (int msl = 1; msl <= nsymp && !rc; ++msl) {
                    int yr = 1, ysaut = 0;
                    const int yrs = msk == msl;
                    const int ypqrs = msi == msk && msj == msl;
                    int ncpqkl = 0, kksvn = 0, llsvn = 0;
                    S.nh = 0;
                    for (int ii = 1; ii <= nsym; ++ii)
                        for (int jj = 1; jj <= ii; ++jj) {
                            int64_t mono1 = (int64_t)nbp[ii - 1] * nbp[jj - 1];
                            if (ypq) mono1 = (mono1 + nbp[jj - 1]) / 2;
                            const int yij = ii == jj;
                            const int nij = ii * (ii - 1) / 2 + jj;
                            const int ijs = ITS(ii, jj);
                            int64_t mij = (int64_t)nbo[ii - 1] * nbo[jj - 1];
                            if (yij) mij = (mij + nbo[ii - 1]) / 2;
                            for (int kk = 1; kk <= ii; ++kk) {
                                const int llm = kk == ii ? jj : kk;
                                for (int ll = 1; ll <= llm; ++ll) {
                                    int64_t mono2 = (int64_t)nbp[kk - 1] * nbp[ll - 1];
                                    if (yrs) mono2 = (mono2 + nbp[ll - 1]) / 2;
                                    if (isymp[ii - 1] != msi || isymp[jj - 1] != msj ||
                                        isymp[kk - 1] != msk || isymp[ll - 1] != msl)
                                        continue;
                                    if (mono1 <= 0 || mono2 <= 0) continue;
                                    const int ykl = kk == ll;
                                    int64_t nijkl = mono1 * mono2;
                                    if (ypqrs) nijkl = (mono1 + nijkl) / 2;
                                    int yijkl = ii == kk && jj == ll;
                                    const int kls = ITS(kk, ll);
                                    if (ITS(kls, ijs) != 1) continue;
                                    const int nkl = kk * (kk - 1) / 2 + ll;
                                    int64_t mkl = (int64_t)nbo[kk - 1] * nbo[ll - 1];
                                    if (ykl) mkl = (mkl + nbo[kk - 1]) / 2;
                                    const int ijkls = nij * (nij - 1) / 2 + nkl;
                                    int64_t wijkl = mkl * mij;
                                    if (wijkl <= 0) continue;
                                    /* itijkl.f:121-200: symmetry transformations */
                                    int skip = 0;
                                    for (int t = 1; t <= ntrsy && !skip; ++t) {
                                        int isig = 1;
                                        int iit = ISYDEG(ii, t), jjt, kkt, llt;
                                        if (iit == 0) continue;
                                        if (iit < 0) { iit = -iit; isig = -1; }
                                        jjt = ISYDEG(jj, t);
                                        if (jjt == 0) continue;
                                        if (jjt < 0) { jjt = -jjt; isig = -isig; }
                                        kkt = ISYDEG(kk, t);
                                        if (kkt == 0) continue;
                                        if (kkt < 0) { kkt = -kkt; isig = -isig; }
                                        llt = ISYDEG(ll, t);
                                        if (llt == 0) continue;
                                        if (llt < 0) { llt = -llt; isig = -isig; }
                                        int nklst = 2;                       /* 180 */
                                        if (iit == kk && jjt == ll && kkt == ii && llt == jj) yijkl = 1;
                                        if (iit < jjt) { int x = iit; iit = jjt; jjt = x; nklst += 1; }
                                        if (kkt < llt) { int x = kkt; kkt = llt; llt = x; nklst += 2; }
                                        if (iit < kkt || (iit == kkt && llt > jjt)) {    /* 182-184 */
                                            int x = iit; iit = kkt; kkt = x;
                                            x = jjt; jjt = llt; llt = x;
                                            nklst += 4;
                                        }
                                        const int ijt = iit * (iit - 1) / 2 + jjt;   /* 185 */
                                        const int klt = kkt * (kkt - 1) / 2 + llt;
                                        if (ijt > nij || (ijt == nij && klt >= nkl)) {  /* 186 */
                                            if (ITS(ITS(iit, jjt), ITS(kkt, llt)) != 1) skip = 1;
                                            continue;
                                        }
                                        const int ijklt = ijt * (ijt - 1) / 2 + klt;     /* 187 */
                                        if (ijkls > nsyt) { rc = -1; break; }
                                        if (ktypsy[ijklt - 1] == 10 || ktypsy[ijklt - 1] == 1) {
                                            if (nklst == 5) nklst = 10;
                                            if (nklst != 10) nklst = 1;
                                        }
                                        ktypsy[ijkls - 1] = kLtabs[(nklst - 1) + 10 * (ktypsy[ijklt - 1] - 1)];
                                        nad[ijkls - 1] = isig * nad[ijklt - 1];
                                        skip = 1;
                                    }
                                    if (rc) break;
                                    if (skip) continue;
                                    if (yr) {
                                        memset(blk, 0, sizeof(double) * (size_t)nijkl);
                                        const int64_t last = sym_transd(&S, ii, jj, kk, ll, blk);
                                        kksvn = kk;
                                        llsvn = ll;
                                        yr = 0;
                                        ysaut = last == 0;
                                    }
                                    if (ysaut) continue;
                                    if (npair / 2 >= pair_cap || ijkls > nsyt) { rc = -1; break; }
                                    npair += 2;
                                    ipair[npair - 2] = nij;
                                    ipair[npair - 1] = nkl;
                                    ktypsy[ijkls - 1] = yijkl ? 1 : 2;
                                    nad[ijkls - 1] = npair / 2;
                                    sym_tpqkld(&S, nbp[ll - 1], nbp[kk - 1], nbp[jj - 1], nbp[ii - 1],
                                               nbo[jj - 1], nbo[ii - 1], yrs, ypq, ypqrs, yij,
                                               ind[jj - 1], ind[ii - 1], blk, tab, tac);
                                    if (ncpqkl >= 4096) { rc = -1; break; }
                                    rec[ncpqkl].npqkl = mono2 * mij;
                                    rec[ncpqkl].mi = nbo[ll - 1]; rec[ncpqkl].mj = nbo[kk - 1];
                                    rec[ncpqkl].mk = nbo[jj - 1]; rec[ncpqkl].ml = nbo[ii - 1];
                                    rec[ncpqkl].ic = ind[ll - 1]; rec[ncpqkl].jc = ind[kk - 1];
                                    rec[ncpqkl].ykl = ykl; rec[ncpqkl].yij = yij;
                                    rec[ncpqkl].yijkl = yijkl;
                                    ++ncpqkl;
                                }
                                if (rc) break;
                            }
                            if (rc) break;
                        }
                    if (rc || S.oom) { rc = -1; break; }
                    /* itijkl.f:248-285: read the unit-40 stream back block by block */
                    int64_t off = 0;
                    for (int t = 0; t < ncpqkl; ++t) {
                        sym_tijkld(&S, rec[t].mi, rec[t].mj, rec[t].mk, rec[t].ml, rec[t].ic,
                                   rec[t].jc, rec[t].ykl, rec[t].yij, rec[t].yijkl,
                                   nbp[llsvn - 1], nbp[kksvn - 1], yrs, S.h + off, tab, tac);
                        off += rec[t].npqkl;
                    }
                }
        }
#undef ITS
#undef ISYDEG
    *npaire = npair / 2;
    *n_out = S.nout;
    if (S.oom) rc = -1;
    free(blk); free(tab); free(tac); free(rec); free(S.h);
    return rc;
}
