/*
 * cipsi_ijkl.h -- C ABI of the MI355X (gfx950) ijkl AO->MO integral transform.
 *
 * Drop-in for the `ijkl` step of the CIPSI chain (C1 case): same inputs
 * (symmetry-packed AO integrals or the pshf label stream, MO coefficients) and the
 * same packed (pq|rs) output order as the reference.  Plain C types only, callable
 * from Fortran through ISO_C_BINDING (module cipsi_ijkl_mod,
 * cipsi_amd/fortran/cipsi_ijkl_mod.f90).
 *
 * Layouts (0-based; pair0(a,b) = a(a+1)/2 + b for a >= b, P_A = N(N+1)/2, P_M = M(M+1)/2):
 *   ao_packed  P_A(P_A+1)/2 doubles, ao_packed[pair0(pair0(i,j), pair0(k,l))] = (ij|kl)
 *              for i>=j, k>=l, ij>=kl -- the block transd builds (transd.f:369-372).
 *   c          column-major N_AO x N_MO, leading dimension ldc >= N_AO: the active MO
 *              columns /om/ c (vector.f:38-41, after rotat2.f:50-60 frozen compaction).
 *   h_half     P_A x P_M doubles, h_half[pq*P_M + pair0(l,k)] (tpqkld.f:91-102).
 *   mo_packed  P_M(P_M+1)/2 doubles, mo_packed[pair0(kl, pair0(j,i))] = (ij|kl), i<=j,
 *              ij<=kl, k<=l: the order tijkld streams to f50 (tijkld.f:31-79).
 *
 * Every entry point returns CIPSI_OK (0) or a negative status; none calls exit().
 * Host-pointer entry points are synchronous and retain no pointer.  *_dev entry
 * points take device pointers, are asynchronous on the context stream and are
 * completed by cipsi_ijkl_synchronize() (except the label-stream ones, which read a
 * bad-label flag back and synchronize before returning).  After one warm-up call of
 * the same size, precision and plan, the *_dev transforms of a packed block launch only
 * kernels and may be captured in a graph; the graph stays valid until
 * cipsi_ijkl_release_workspace() or cipsi_ijkl_destroy() (a workspace that must grow
 * retires the old buffers instead of freeing them).  A context is not thread-safe; use one
 * context per host thread (the reference runs one process per geometry,
 * Autocip_13/autocip13.f90:525-567).
 */
#ifndef CIPSI_IJKL_H
#define CIPSI_IJKL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes; the reference reports these conditions with `stop`
   (ijkl.f:118-120, ijkl.f:304-316, itijkl.f:230-234, rotat2.f:347-380) */
#define CIPSI_OK                  0
#define CIPSI_ERR_ARG            -1   /* bad dimension, null pointer, ldc < n_ao, n_mo > n_ao */
#define CIPSI_ERR_LABEL          -2   /* label stream: n_ao > 255 or a label outside the block */
#define CIPSI_ERR_NOMEM          -3   /* device allocation failed */
#define CIPSI_ERR_HIP            -4   /* other HIP runtime failure */
#define CIPSI_ERR_NONFINITE      -5   /* NaN or Inf in an input */
#define CIPSI_ERR_NOT_ORTHONORMAL -6  /* |C^T S C - 1| above tolerance */
#define CIPSI_ERR_UNSUPPORTED    -7   /* n_ao > 256, or a precision / slice count / tile variant not built */
#define CIPSI_ERR_NODEVICE       -8   /* no usable gfx950 device */

/* precision argument */
#define CIPSI_PREC_FP64        0      /* FP64 MFMA (v_mfma_f64_16x16x4f64) */
#define CIPSI_PREC_SPLIT_FP16  1      /* Ozaki FP16 slices on FP16 MFMA, FP32 accumulation */

typedef struct cipsi_ijkl_ctx cipsi_ijkl_ctx;

/* Context: owns device workspace, a HIP stream and tuned tile plans.
   No counterpart in the reference (its state lives in COMMON blocks,
   itijkl.f:12-22); needed so the library is re-entrant. */
int cipsi_ijkl_create(cipsi_ijkl_ctx **ctx, int device);
int cipsi_ijkl_destroy(cipsi_ijkl_ctx *ctx);
const char *cipsi_ijkl_strerror(int status);
const char *cipsi_ijkl_last_error(const cipsi_ijkl_ctx *ctx);
/* Run subsequent work on an external hipStream_t, taken as given: NULL is the HIP
   null stream (ordered with every blocking stream of the device). */
int cipsi_ijkl_set_stream(cipsi_ijkl_ctx *ctx, void *hip_stream);
/* Return to the context's own (non-blocking) stream. */
int cipsi_ijkl_reset_stream(cipsi_ijkl_ctx *ctx);
int cipsi_ijkl_synchronize(cipsi_ijkl_ctx *ctx);
/* Device bytes the context allocates for one transform of this size. */
int64_t cipsi_ijkl_workspace_bytes(int n_ao, int n_mo, int precision, int with_labels);
/* Release the cached workspace. */
int cipsi_ijkl_release_workspace(cipsi_ijkl_ctx *ctx);

/* Whole transform of a packed AO block: replaces `call itijkl(bijkl,bijkl,ydp)`
   (ijkl.f:318) for C1 with YDP=T, i.e. tpqkld (itijkl.f:214-220) followed by tijkld
   (itijkl.f:253-285), without the unit-40 scratch file.  err_est (may be NULL)
   receives an a-priori bound on max |error| of the output.  The copy of mo to the
   caller runs while half 2 still computes (FP64); after an error return the contents
   of mo are unspecified. */
int cipsi_ijkl_transform_packed(cipsi_ijkl_ctx *ctx, int n_ao, int n_mo,
                                const double *ao_packed, const double *c, int ldc,
                                double *mo_packed, int precision, int n_slices,
                                double *err_est);

/* Same from the pshf/cval label stream (itijkl.f:57-68 + transd.f): xx and labels
   hold n_entries values and 32-bit labels i<<24|j<<16|k<<8|l (iword_func.f:2-52,
   1-based indices, values pre-divided as pshf/qout.f:73-77 does); the first zero
   label ends the stream; a later duplicate overwrites an earlier one.  Labels are
   expected in pshf's canonical order i>=j, k>=l, ij>=kl (pshf/qout.f:52-69).  Like the
   reference's C1 route, which stores the label as read (transd.f:356,369-372: the
   canonicalisation there applies to symmetry images only), a non-canonical label is
   stored at pair0(pair0(i,j), pair0(k,l)) as given, doubled for i==j, k==l and
   (i,j)==(k,l) of its own fields (transd.f:211-218); CIPSI_ERR_LABEL is returned only
   for a field outside 1..n_ao. */
int cipsi_ijkl_transform_labels(cipsi_ijkl_ctx *ctx, int n_ao, int n_mo, const double *xx,
                                const int32_t *labels, int64_t n_entries, const double *c,
                                int ldc, double *mo_packed, int precision, int n_slices,
                                double *err_est);

/* Device-pointer forms of the two entry points above. */
int cipsi_ijkl_transform_packed_dev(cipsi_ijkl_ctx *ctx, int n_ao, int n_mo,
                                    const double *d_ao_packed, const double *d_c, int ldc,
                                    double *d_mo_packed, int precision, int n_slices);
int cipsi_ijkl_transform_labels_dev(cipsi_ijkl_ctx *ctx, int n_ao, int n_mo, const double *d_xx,
                                    const int32_t *d_labels, int64_t n_entries,
                                    const double *d_c, int ldc, double *d_mo_packed,
                                    int precision, int n_slices);

/* The same FP64 transform as cipsi_ijkl_transform_packed_dev (same kernels, tile plan and
   bits) for callers that cannot allocate device memory: code under graph capture, tight
   serving loops, harnesses that forbid hipMalloc.  No counterpart in the reference.  It
   calls no hipMalloc / hipFree: its workspace is a static 22.3 MB device array, one copy
   per GPU, held until the process ends, which limits it to n_ao <= 48
   (CIPSI_ERR_UNSUPPORTED above).  FP64 only.  Asynchronous on the context stream.  Every
   call on one GPU uses that GPU's single copy, so calls that overlap on one GPU (two
   streams, two threads, two contexts) race: serialise them. */
int cipsi_ijkl_transform_packed_noalloc_dev(cipsi_ijkl_ctx *ctx, int n_ao, int n_mo,
                                            const double *d_ao_packed, const double *d_c,
                                            int ldc, double *d_mo_packed);

/* ---- Point-group symmetry (nsym > 1) ---------------------------------------------
 * The state itijkl reads from COMMON once ijkl.f has run rondof, otodon, rotat2 and
 * vector (ijkl.f:186-350): the in-process seam `call itijkl(bijkl,bijkl,ydp)`
 * (ijkl.f:405) for molecules with symmetry.  All indices 1-based as in the reference;
 * Fortran arrays column-major.
 *   n_ao, nsym, nt (group operations), nshell
 *   its      nsym x nsym product table its(i,j)                 (blank COMMON)
 *   isymp    class of each irrep, non-decreasing (rotat2.f:114-157)
 *   nbp,nbo  AOs / MOs of each irrep; ind: nsym+1 offsets into c (vector.f:22-46)
 *   ivecsy   nsym x n_ao, ivecsy[(irrep-1)*n_ao + ao-1]: index of the AO in
 *            its irrep's subset, 0 if absent (vector.f:26-28)
 *   c        /om/ c: per irrep, nbp x nbo column-major from ind(irrep)
 *   ishell, ktyp, numc   per AO: shell, shell type, component (rondof.f:77-82)
 *   kloc     per shell: first AO; newsh: nshell x nt image shells (rondof.f:30-38)
 *   iptr, idtr, iftr     3 x nt, 6 x nt, 10 x nt diagonal signs (rondof.f:100-118)
 *   ntrsy, isydeg  the symmetry transformations of the linear groups (blank COMMON,
 *            ijkl.f:141-149): nsym x ntrsy signed irrep maps, column-major; ntrsy = 0
 *            (isydeg unused) for the other groups or with ntrans = 0 in the namelist.
 *            itijkl skips every irrep quartet a transformation maps onto an earlier one
 *            (itijkl.f:121-200).
 * Every AO must belong to the irreps of one class only (true for diatomics in C2v, D2h,
 * C-inf-v and D-inf-h; not for a linear molecule with an atom at the centre). */
typedef struct {
    int n_ao, nsym, nt, nshell;
    const int *its, *isymp, *nbp, *nbo, *ind, *ivecsy;
    const double *c;
    const int *ishell, *kloc, *ktyp, *numc, *newsh, *iptr, *idtr, *iftr;
    int ntrsy;
    const int *isydeg;
} cipsi_ijkl_sym;

/* otodon.f:163-317: irreps, product table its (nsym x nsym, column-major) and characters
   jsym (nsym x nt, irrep x operation) of a point group without degeneracies, from the
   info file's group name (8 characters, blank padded, either case) and naxis: C1, CS,
   CI, CN, CNV, CNH, DN, DNH.  CINFV / DINFH (their tables depend on the basis: use
   cipsi_ijkl_sym_linear) and an axis order other than 2 give CIPSI_ERR_UNSUPPORTED; an
   unknown name CIPSI_ERR_ARG. */
int cipsi_ijkl_sym_group(const char *group, int naxis, int nt, int *nsym, int *its, int *jsym);

/* otodon.f:36-144: the linear groups CINFV and DINFH, as pshf hands them over (the info
   file names the linear group and carries the operations of C1, or of Cs = (E, sigma_h)
   for DINFH, ptgrp.f).  ijkl classifies the AOs by their angular character about the
   axis: sigma, pi_x, pi_y, delta_x, delta_y (phi_x, phi_y with f shells), times g / u for
   DINFH, whose symmetry orbitals are an AO plus or minus its image on the other atom.
   Inputs per shell: ktype, kloc (1-based first AO), zsh (z coordinate of the shell's
   atom), and for DINFH newsh (nshell x nt, column-major, nt >= 2: column 2 is the image
   under sigma_h).  Outputs: nsym (5 or 7; 10 or 14), its (nsym x nsym, column-major,
   room for 14 x 14), ntrans (0-2: one symmetry transformation when p shells are
   present, two with d shells) and isydeg (nsym x ntrans, column-major, room for 14 x 2),
   so (nsym x n_ao, room for 14 n_ao): otodon's signed symmetry-orbital numbers, the
   input of cipsi_ijkl_sym_prepare_so. */
int cipsi_ijkl_sym_linear(const char *group, int n_ao, int nshell, const int *ktype,
                          const int *kloc, const double *zsh, const int *newsh, int nt,
                          int *nsym, int *its, int *ntrans, int *isydeg, int *so);

/* The MO preparation ijkl does before itijkl for nsym > 1, host FP64 (no device work):
   otodon.f:320-353 (AOs of each irrep's symmetry orbitals), rotat2.f:50-423 (frozen
   compaction with ygel[0..nvecf), classes of irreps, degenerate levels from the MO
   diagonal of h, projection onto the irreps, normalisation with S and orthogonality
   check, MOs reordered occupied-by-irrep then virtual-by-irrep) and vector.f:17-46.
   Inputs as the info file holds them: basis tables per shell (ktype, kloc, newsh
   nshell x nt), the diagonal signs iptr / idtr / iftr, S and h packed (s_packed NULL: S
   = 1, the reference's yatom), C square column-major (record 7).  Outputs: c_out (the
   coefficients written to ijom, n_ao x n_ao), itsyv (irrep of each of the nvecf MOs),
   isymp / nsymp (classes), and the COMMON state of cipsi_ijkl_sym: ivecsy, nbp, nbo, ind
   and c_packed (at most n_ao x n_ao doubles).  Where the reference stops with an error
   tag (rotat2.f:261-289,347-354,373-380) this returns CIPSI_ERR_NOT_ORTHONORMAL with the
   tag (2: an MO overlaps too few irreps, 1: too many / not orthogonal, -1: zero norm) in
   *tag.  Sums run in the reference's order without fused multiply-add, so c_out equals
   the reference's coefficients bit for bit. */
int cipsi_ijkl_sym_prepare(int n_ao, int nsym, const int *its, const int *jsym, int nt,
                           int nshell, const int *ktype, const int *kloc, const int *newsh,
                           const int *iptr, const int *idtr, const int *iftr,
                           const double *s_packed, const double *h_packed,
                           const double *c_square, int nvecf, const int *ygel, int noc,
                           double trec, double tdeg, double *c_out, int *itsyv, int *isymp,
                           int *nsymp, int *ivecsy, int *nbp, int *nbo, int *ind,
                           double *c_packed, int *tag);

/* cipsi_ijkl_sym_prepare after otodon has run elsewhere: rotat2.f:50-423 and
   vector.f:17-46 for the symmetry-orbital table so (nsym x n_ao, the signed numbers
   otodon leaves in ivecsy), the route of the linear groups (cipsi_ijkl_sym_linear).
   nsym <= 14; the other arguments and the outputs as cipsi_ijkl_sym_prepare. */
int cipsi_ijkl_sym_prepare_so(int n_ao, int nsym, const int *so, const double *s_packed,
                              const double *h_packed, const double *c_square, int nvecf,
                              const int *ygel, int noc, double trec, double tdeg,
                              double *c_out, int *itsyv, int *isymp, int *nsymp, int *ivecsy,
                              int *nbp, int *nbo, int *ind, double *c_packed, int *tag);

/* itijkl.f:42-298 for nsym > 1: the pshf skeleton label stream (transd.f regenerates the
   symmetry images, transd.f:125-244) -> the f50 MO stream of every symmetry-allowed
   block, in itijkl's order (class quartets, then irrep quartets), each block in
   tijkld's order (tijkld.f:31-79), and the block directory written to f25
   (itijkl.f:208-213,291-296): npaire blocks, ipair[2b], ipair[2b+1] = (nij, nkl),
   nad / ktypsy over nsyt = min(2000, P(P(nsym))) entries.  mo_count receives the
   stream length; with mo_stream == NULL only the sizes and directory are computed (the
   label stream is still unpacked: which class quartets hold integrals decides the
   directory).  With symmetry transformations (sym->ntrsy > 0) a block equivalent to an
   earlier one is left out of the stream; its directory entry points at that block (nad
   negative when the transformation changes the sign, ktypsy the index permutation,
   itijkl.f:121-200), and a block the transformations map onto itself with its two pairs
   exchanged is stored as a triangle (ktypsy 1).  A directory beyond nsyt entries (D-inf-h
   with f shells) gives CIPSI_ERR_UNSUPPORTED.  Host pointers, synchronous.  FP64 MFMA. */
int cipsi_ijkl_sym_transform_labels(cipsi_ijkl_ctx *ctx, const cipsi_ijkl_sym *sym,
                                    const double *xx, const int32_t *labels, int64_t n_entries,
                                    double *mo_stream, int64_t mo_capacity, int64_t *mo_count,
                                    int *ipair, int max_pairs, int *npaire, int16_t *nad,
                                    int16_t *ktypsy);

/* transd.f (C1 route, transd.f:85-97,210-218,356-373): label stream -> packed AO block
   on the device.  Returns CIPSI_ERR_LABEL (after synchronising) for a bad label. */
int cipsi_ijkl_unpack_labels_dev(cipsi_ijkl_ctx *ctx, int n_ao, const double *d_xx,
                                 const int32_t *d_labels, int64_t n_entries,
                                 double *d_ao_packed);

/* tpqkld.f:25-104: first half-transform, packed AO block -> h_half (P_A x P_M). */
int cipsi_ijkl_half1_dev(cipsi_ijkl_ctx *ctx, int n_ao, int n_mo, const double *d_ao_packed,
                         const double *d_c, int ldc, double *d_h_half, int precision,
                         int n_slices);

/* tijkld.f:28-88: second half-transform, h_half (P_A x P_M) -> packed MO stream. */
int cipsi_ijkl_half2_dev(cipsi_ijkl_ctx *ctx, int n_ao, int n_mo, const double *d_h_half,
                         const double *d_c, int ldc, double *d_mo_packed, int precision,
                         int n_slices);

/* tmono.f:39-54: h_MO = C^T h C, h packed N(N+1)/2 -> h_mo packed M(M+1)/2 (host). */
int cipsi_ijkl_tmono(cipsi_ijkl_ctx *ctx, int n_ao, int n_mo, const double *h_packed,
                     const double *c, int ldc, double *h_mo_packed);

/* Energy and its gradient with respect to the MO coefficients (the adjoint of the
 * transform).  The reference has no such entry point: its consumers (fock.f:80-101,
 * and the CI energy assembled downstream) evaluate the same contraction
 *   E(C) = sum_pq h^MO_pq gamma_pq + 1/2 sum_pqrs Gamma_pqrs (pq|rs)_MO
 * on the tmono / tijkld outputs; this computes E and dE/dC on the device from the
 * AO inputs, recomputing the half-transformed integrals instead of storing them.
 *   gamma        M x M symmetric one-particle density (row-major)
 *   Gamma_packed P_M(P_M+1)/2 pair density, packed like mo_packed:
 *                Gamma_pqrs = Gamma_packed[pair0(max(a,b),min(a,b))], a = pair(p,q),
 *                b = pair(r,s), for every index order
 *   grad         N_AO x N_MO column-major (leading dimension N_AO); may be NULL
 *                (energy only).
 * FP64 throughout; one rocBLAS DGEMM (P_A x P_M x P_M) forms the pair-density
 * contraction.  Workspace adds P_M(P_M+1)/2 + P_M^2 doubles to the transform's. */
int cipsi_ijkl_energy_grad(cipsi_ijkl_ctx *ctx, int n_ao, int n_mo, const double *ao_packed,
                           const double *h_packed, const double *c, int ldc,
                           const double *gamma, const double *Gamma_packed, double *energy,
                           double *grad);
int cipsi_ijkl_energy_grad_dev(cipsi_ijkl_ctx *ctx, int n_ao, int n_mo,
                               const double *d_ao_packed, const double *d_h_packed,
                               const double *d_c, int ldc, const double *d_gamma,
                               const double *d_Gamma_packed, double *d_energy, double *d_grad);

/* Tile autotuner: times every built tile variant of the two half-transform kernels
   on device-resident synthetic data of this size and keeps the fastest per half in the
   context (used by every later call with the same n_ao, n_mo, precision).  best_half1 /
   best_half2 (may be NULL) receive the chosen variant indices.  No counterpart in the
   reference; tile shapes are specific to this implementation. */
int cipsi_ijkl_autotune(cipsi_ijkl_ctx *ctx, int n_ao, int n_mo, int precision,
                        int *best_half1, int *best_half2);
/* Set / read the plan (variant indices) used for a size; -1 leaves a half unchanged.
   A 12-wave variant is built only for N_AO <= 192: set_plan (and
   cipsi_ijkl_batch_set_class_plan for a row class above 12 blocks of 16) returns
   CIPSI_ERR_UNSUPPORTED for it, and the autotuners skip it there. */
int cipsi_ijkl_set_plan(cipsi_ijkl_ctx *ctx, int n_ao, int n_mo, int precision, int var_half1,
                        int var_half2);
int cipsi_ijkl_get_plan(cipsi_ijkl_ctx *ctx, int n_ao, int n_mo, int precision, int *var_half1,
                        int *var_half2);
/* Number of built tile variants, and the shape of one (waves per work-group, panel depth,
   work-groups per CU hint). */
int cipsi_ijkl_num_variants(void);
int cipsi_ijkl_variant_info(int variant, int *waves, int *panel_depth, int *min_blocks);

/* Per-kernel GPU timing with HIP events recorded on the context stream around each
   launch of the named stage ("expand", "half1", "transpose", "half2", "pad_c",
   "unpack").  For benchmarking; no counterpart in the reference (whose timer calls
   are commented out, itijkl.f:202-203,221-222). */
int cipsi_ijkl_enable_timing(cipsi_ijkl_ctx *ctx, int enable);
int cipsi_ijkl_read_timing(cipsi_ijkl_ctx *ctx, const char *stage, double *total_ms,
                           int64_t *launches);
int cipsi_ijkl_reset_timing(cipsi_ijkl_ctx *ctx);

/* ---- Batched scan of geometries --------------------------------------------------
 * The reference transforms one geometry per ijkl process; its scan driver starts one
 * process per geometry (Autocip_13/autocip13.f90:525-567, each reaching ijkl.f:318).
 * A batch plan transforms all geometries of a scan, each with its own N_AO, N_MO, AO
 * block and C, in one hipGraph launch of grouped kernels; the per-geometry
 * results are those of cipsi_ijkl_transform_packed_dev.
 *   create      sizes only; allocates the device workspace of every geometry
 *               (cipsi_ijkl_batch_workspace_bytes) and per-geometry argument tables.
 *   autotune    times every tile variant of both halves for each row class of the
 *               plan on the given data and keeps the fastest (leaves the outputs of a
 *               full run in d_mo).
 *   run         d_ao[g], d_c[g] (ldc[g] >= n_ao[g]) and d_mo[g] are device pointers,
 *               held in host arrays of n_geom entries; asynchronous on the context
 *               stream.  use_graph != 0 launches the captured graph (captured on the
 *               first such run and after a plan change), 0 launches the same kernels
 *               directly.  Runs of one plan must stay on one stream.
 *   class_info  row class k of the plan: row blocks (padded N_AO / 16 bound), number
 *               of geometries and the tile variants of the two halves. */
typedef struct cipsi_ijkl_batch cipsi_ijkl_batch;
int64_t cipsi_ijkl_batch_workspace_bytes(int n_geom, const int *n_ao, const int *n_mo);
int cipsi_ijkl_batch_create(cipsi_ijkl_ctx *ctx, int n_geom, const int *n_ao, const int *n_mo,
                            cipsi_ijkl_batch **plan);
int cipsi_ijkl_batch_destroy(cipsi_ijkl_ctx *ctx, cipsi_ijkl_batch *plan);
int cipsi_ijkl_batch_autotune(cipsi_ijkl_ctx *ctx, cipsi_ijkl_batch *plan,
                              const double *const *d_ao, const double *const *d_c,
                              const int *ldc, double *const *d_mo);
int cipsi_ijkl_batch_run(cipsi_ijkl_ctx *ctx, cipsi_ijkl_batch *plan, const double *const *d_ao,
                         const double *const *d_c, const int *ldc, double *const *d_mo,
                         int use_graph);
int cipsi_ijkl_batch_num_classes(const cipsi_ijkl_batch *plan);
/* One call per scan, the batch entry point of the drop-in proposal (no counterpart in
   the reference, whose scan starts one ijkl process per geometry): device pointers as in
   cipsi_ijkl_batch_run, FP64 only (precision 0, n_slices 0; split-FP16 gives
   CIPSI_ERR_UNSUPPORTED).  The context keeps one plan, for its last size list: a new
   size list builds and autotunes a plan on this call's data (seconds for 64 geometries)
   and synchronizes; the same size list again replays the captured graph, asynchronous
   on the context stream.  cipsi_ijkl_release_workspace and cipsi_ijkl_destroy free the
   plan. */
int cipsi_ijkl_transform_batch_dev(cipsi_ijkl_ctx *ctx, int n_geom, const int *n_ao,
                                   const int *n_mo, const double *const *d_ao,
                                   const double *const *d_c, const int *ldc,
                                   double *const *d_mo, int precision, int n_slices);
int cipsi_ijkl_batch_class_info(const cipsi_ijkl_batch *plan, int k, int *row_blocks,
                                int *n_geom, int *var_half1, int *var_half2);
int cipsi_ijkl_batch_set_class_plan(cipsi_ijkl_ctx *ctx, cipsi_ijkl_batch *plan, int k,
                                    int var_half1, int var_half2);

/* ---- Integral consumers of the packed MO stream ----------------------------------
 * gather: out[q] = ai(i,j,k,l) for the 1-based MO indices of labels[q] (the label codec
 *   of ijkl/iword_func.f:2-52, i<<24 | j<<16 | k<<8 | l); the lookup of fock/ai.f:12-35
 *   for C1, pair indices with the larger pair first, for any index order.  Device
 *   pointers; n_mo <= 255.  A label with a field outside 1..n_mo gives NaN and the call
 *   returns CIPSI_ERR_LABEL.  Synchronous at return (it reads back that count).
 * fock: the closed-shell electronic energy (without the nuclear term), the MO Fock
 *   operator with |F| < 1e-7 set to zero, and the J = (ii|jj) and K = (ij|ij) matrices of
 *   CIPSI_sources/fock/fock.f:80-127, for noc doubly occupied MOs (the first noc).
 *   h_mo_packed is tmono's output; fock, j, k are packed like it (i >= j).  Sums run in
 *   fock.f's order without fused multiply-adds, so on equal MO integrals the results
 *   equal those of fock.f built without FMA.  The _dev form is asynchronous on the
 *   context stream; the host form copies in and out and is synchronous. */
int cipsi_ijkl_gather_dev(cipsi_ijkl_ctx *ctx, int n_mo, const double *d_mo_packed,
                          const int32_t *d_labels, int64_t n, double *d_out);
int cipsi_ijkl_fock(cipsi_ijkl_ctx *ctx, int n_mo, int noc, const double *mo_packed,
                    const double *h_mo_packed, double *energy, double *fock, double *jint,
                    double *kint);
int cipsi_ijkl_fock_dev(cipsi_ijkl_ctx *ctx, int n_mo, int noc, const double *d_mo_packed,
                        const double *d_h_mo_packed, double *d_energy, double *d_fock,
                        double *d_j, double *d_k);

/* Slater-Condon elements of cip between determinants held, as cip holds them (common
 * /spo/), as excitations of the closed-shell reference; all numbers 1-based:
 *   determinant d (0-based): ne[d] <= 10 excitations, holes trou[nd[d] + t] and particles
 *   part[nd[d] + t] (t < ne[d]) as spin-orbital numbers 1..2 n_mo; spin-orbital s has
 *   orbital iorb[s-1] (1..n_mo) and spin ispin[s-1] (1 or 2).  n_exc = length of trou
 *   and part.  fock, jint, kint: the packed F, J, K of cipsi_ijkl_fock (what cip reads
 *   from fock's unit 40, cip/ijkf.f:26-36).
 * hdig: out[d] = E(D_d) - E(reference) (cip/hdig.f:40-80, fdiag = diagonal of F).
 * hntd: out[q] = <D_a|H|D_b> for pairs[2q] = a, pairs[2q+1] = b (0-based), zero when the
 *   determinants are identical or differ by more than two spin-orbitals
 *   (cip/hntd.f:36-168), with ai() on the packed MO stream.  cip keeps that stream in
 *   REAL*4 (cip/ai.f); this reads the FP64 stream.
 * Device pointers; the arrays are validated first (CIPSI_ERR_ARG on a number out of
 * range or more than 10 excitations).  Synchronous at return.  No reference entry point
 * is replaced: cip calls hdig / hntd one determinant or pair at a time (ic001.f:105,
 * pert1.f:116, diexcit.f:293). */
int cipsi_ijkl_hdig_dev(cipsi_ijkl_ctx *ctx, int n_mo, int n_det, int64_t n_exc,
                        const int *d_ne, const int *d_nd, const int *d_trou, const int *d_part,
                        const int *d_iorb, const int *d_ispin, const double *d_fock,
                        const double *d_jint, const double *d_kint, double *d_out);
int cipsi_ijkl_hntd_dev(cipsi_ijkl_ctx *ctx, int n_mo, int n_det, int64_t n_exc,
                        const int *d_ne, const int *d_nd, const int *d_trou, const int *d_part,
                        const int *d_iorb, const int *d_ispin, const int *d_pairs,
                        int64_t n_pairs, const double *d_mo_packed, const double *d_fock,
                        double *d_out);

#ifdef __cplusplus
}
#endif

#endif /* CIPSI_IJKL_H */
