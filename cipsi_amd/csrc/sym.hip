// Point-group symmetry: the itijkl step for nsym > 1 on the GPU (gfx950).
//
// The reference (itijkl.f:42-298) loops over classes of irreducible representations
// (irreps whose MOs use the same AOs, rotat2.f:114-157) and over the symmetry-allowed
// irrep quartets (A,B|C,D) of each class quartet.  transd fills one AO block per class
// quartet from the symmetry-unique ("skeleton") label stream, regenerating the images of
// every integral under the group (transd.f:125-244); tpqkld and tijkld then transform
// each irrep quartet as a rectangular sandwich with two different C blocks, and the
// outputs follow one another in the f50 stream.
//
// Here:
//   unpack   one pass over the labels writes every image of every integral into a full
//            packed AO block whose AO order is permuted class by class, so that the AOs
//            of every irrep form one contiguous range; last-writer election on 64-bit
//            (label, image) keys keeps transd's sequential semantics for duplicates, and
//            the class quartets that receive an integral are flagged (transd reports an
//            empty class quartet, which itijkl then skips, itijkl.f:199,207,376).
//   expand   packed block -> full symmetric pair matrix (the C1 kernel).
//   half 1   per block (A,B|C,D), per AO pair (p in D, q in C): Z(r, s) = (s r|q p) for
//            s in A, r in B is a sub-rectangle of pair-matrix row (p,q); H = C_B^T Z C_A
//            on FP64 MFMA, the triangle a <= c when A == B (tpqkld.f:28-104).
//   transpose H -> H^T per block (grouped launch).
//   half 2   per MO pair (l,k): Y(p, q) = H[(p,q)][(l,k)]; O = C_D^T Y C_C, the triangle
//            i <= j when C == D, rows of a diagonal block stopping at nij = ilk
//            (tijkld.f:28-79), written at the block's offset of the output stream.
// The linear groups add symmetry transformations (rotations about the axis that map
// irreps onto irreps, itijkl.f:121-200): an irrep quartet that one of them maps onto an
// earlier quartet is not transformed at all, its directory entry points at the earlier
// block, and a quartet mapped onto itself with its pairs exchanged is stored as a
// triangle, as a diagonal block is.  This is host bookkeeping in the block list.
// Classes must be disjoint (every AO in the irreps of one class only), which holds for
// diatomics in C2v, D2h, C-inf-v and D-inf-h; the entry point refuses anything else.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cipsi_ijkl.h"
#include "common.h"
#include "ctx.h"

namespace cipsi {

constexpr int kSymMaxAO = 255;
constexpr int kSymMaxShell = 255;
constexpr int kSymMaxOps = 8;
constexpr int kSymMaxIrreps = 14;         // D-inf-h with f shells
constexpr int kSymMaxClassPairs = kSymMaxIrreps * (kSymMaxIrreps + 1) / 2;
constexpr int kSymMaxClassKeys = kSymMaxClassPairs * (kSymMaxClassPairs + 1) / 2;
constexpr int kSymMaxTrans = 64;          // symmetry transformations (nsymz columns)

// Basis and group tables, 0-based (uploaded once per call).
struct SymTables {
    int n, nt, nshell, nclass;
    int kloc[kSymMaxShell];                    // first AO of each shell
    int newsh[kSymMaxShell * kSymMaxOps];      // image shell, [s * 8 + g]
    int ishell[kSymMaxAO], ktyp[kSymMaxAO], numc[kSymMaxAO];
    int perm[kSymMaxAO];                       // AO -> position in the class-ordered basis
    int cls[kSymMaxAO];                        // AO -> class (0-based), or -1
    int sgn_p[3 * kSymMaxOps], sgn_d[6 * kSymMaxOps], sgn_f[10 * kSymMaxOps];   // [c * 8 + g]
};

// transd's canonical shell quartet (transd.f:134-145 for the label, 167-187 for images):
// a >= b, c >= d, then the pairs ordered with ties broken by the second index.
__device__ inline void canon4(int& a, int& b, int& c, int& d) {
    if (!(a > b)) { const int t = a; a = b; b = t; }
    if (!(c > d)) { const int t = c; c = d; d = t; }
    if (a < c || (a == c && d > b)) {
        int t = a; a = c; c = t;
        t = b; b = d; d = t;
    }
}

// The operations that generate distinct images of a shell quartet (bit g; bit 0 is the
// identity) and the size of its stabiliser, nq4 (transd.f:157-209).
__device__ inline unsigned sym_images(const SymTables& T, int ish, int jsh, int ksh, int lsh,
                                      int* nq4_out) {
    int ic = ish, jc = jsh, kc = ksh, lc = lsh;
    canon4(ic, jc, kc, lc);
    int nq4 = 1, ngen = 0;
    int gen[kSymMaxOps][4];
    unsigned ysy = 1u;
    for (int g = 1; g < T.nt; ++g) {
        int n1 = T.newsh[ic * kSymMaxOps + g], n2 = T.newsh[jc * kSymMaxOps + g];
        int n3 = T.newsh[kc * kSymMaxOps + g], n4 = T.newsh[lc * kSymMaxOps + g];
        canon4(n1, n2, n3, n4);
        if (n1 == ic && n2 == jc && n3 == kc && n4 == lc) { ++nq4; continue; }
        bool dup = false;
        for (int x = 0; x < ngen; ++x)
            dup = dup || (gen[x][0] == n1 && gen[x][1] == n2 && gen[x][2] == n3 && gen[x][3] == n4);
        if (dup) continue;
        gen[ngen][0] = n1; gen[ngen][1] = n2; gen[ngen][2] = n3; gen[ngen][3] = n4;
        ++ngen;
        ysy |= 1u << g;
    }
    *nq4_out = nq4;
    return ysy;
}

__device__ inline int comp_sign(const SymTables& T, int ao, int g) {
    const int kt = T.ktyp[ao], c = T.numc[ao];
    if (kt == 1) return 1;
    if (kt == 2) return T.sgn_p[c * kSymMaxOps + g];
    if (kt == 3) return T.sgn_d[c * kSymMaxOps + g];
    return T.sgn_f[c * kSymMaxOps + g];
}

__device__ inline int image_ao(const SymTables& T, int ao, int g) {
    if (g == 0) return ao;
    return T.kloc[T.newsh[T.ishell[ao] * kSymMaxOps + g]] + T.numc[ao];
}

// sc[0] = first zero label (the stream end), sc[1] = first label with a field outside
// 1..n; both start at n.
__global__ void k_sym_scan_init(unsigned long long* sc, int64_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sc[0] = (unsigned long long)n;
        sc[1] = (unsigned long long)n;
    }
}

__global__ __launch_bounds__(256) void k_sym_scan(const int32_t* __restrict__ labels, int64_t n,
                                                  int n_ao, unsigned long long* sc) {
    unsigned long long zmin = ~0ull, bmin = ~0ull;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        const uint32_t u = (uint32_t)labels[t];
        if (u == 0u) { zmin = min(zmin, (unsigned long long)t); continue; }
        const int i = (int)(u >> 24), j = (int)((u >> 16) & 255u), k = (int)((u >> 8) & 255u),
                  l = (int)(u & 255u);
        if (i < 1 || j < 1 || k < 1 || l < 1 || i > n_ao || j > n_ao || k > n_ao || l > n_ao)
            bmin = min(bmin, (unsigned long long)t);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        zmin = min(zmin, (unsigned long long)__shfl_xor(zmin, o));
        bmin = min(bmin, (unsigned long long)__shfl_xor(bmin, o));
    }
    if ((threadIdx.x & 63) == 0) {
        if (zmin < sc[0]) atomicMin(&sc[0], zmin);
        if (bmin < sc[1]) atomicMin(&sc[1], bmin);
    }
}

__device__ inline int class_key(int a, int b, int c, int d) {
    const int p1 = a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a;
    const int p2 = c >= d ? c * (c + 1) / 2 + d : d * (d + 1) / 2 + c;
    return p1 >= p2 ? p1 * (p1 + 1) / 2 + p2 : p2 * (p2 + 1) / 2 + p1;
}

// Every image of every label before the stream end claims its position in the
// class-ordered packed block with key 8 * label + image + 1 (the reference writes
// label after label, image after image, so the largest key is the last writer), and
// flags the class quartet it belongs to.
__global__ __launch_bounds__(256) void k_sym_claim(const int32_t* __restrict__ labels,
                                                   const SymTables* __restrict__ Tp,
                                                   const unsigned long long* sc,
                                                   unsigned long long* __restrict__ owner,
                                                   int* __restrict__ touched) {
    const SymTables& T = *Tp;
    const int64_t end = (int64_t)min(sc[0], sc[1]);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < end; t += stride) {
        const uint32_t u = (uint32_t)labels[t];
        const int a0 = (int)(u >> 24) - 1, a1 = (int)((u >> 16) & 255u) - 1;
        const int a2 = (int)((u >> 8) & 255u) - 1, a3 = (int)(u & 255u) - 1;
        int nq4;
        const unsigned ysy = sym_images(T, T.ishell[a0], T.ishell[a1], T.ishell[a2], T.ishell[a3], &nq4);
        for (int g = 0; g < T.nt; ++g) {
            if (!((ysy >> g) & 1u)) continue;
            const int i = image_ao(T, a0, g), j = image_ao(T, a1, g);
            const int k = image_ao(T, a2, g), l = image_ao(T, a3, g);
            const int64_t pij = pairsym(T.perm[i], T.perm[j]), pkl = pairsym(T.perm[k], T.perm[l]);
            atomicMax(owner + pairsym(pij, pkl), (unsigned long long)(8 * t + g) + 1ull);
            const int ci = T.cls[i], cj = T.cls[j], ck = T.cls[k], cl = T.cls[l];
            if (ci >= 0 && cj >= 0 && ck >= 0 && cl >= 0) touched[class_key(ci, cj, ck, cl)] = 1;
        }
    }
}

// Every position of the block: the owning image's value, as transd computes it
// (transd.f:209-237): xx * nq4 / nt, doubled for i == j, k == l and (i,j) == (k,l) of the
// label, negated by the component signs of the image; zero without owner.
__global__ __launch_bounds__(256) void k_sym_resolve(const double* __restrict__ xx,
                                                     const int32_t* __restrict__ labels,
                                                     const SymTables* __restrict__ Tp,
                                                     const unsigned long long* __restrict__ owner,
                                                     int64_t nblk, double* __restrict__ out) {
    const SymTables& T = *Tp;
    const double dfnt = 1.0 / (double)T.nt;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nblk; p += stride) {
        const unsigned long long w = owner[p];
        if (w == 0ull) { out[p] = 0.0; continue; }
        const int64_t t = (int64_t)((w - 1ull) >> 3);
        const int g = (int)((w - 1ull) & 7ull);
        const uint32_t u = (uint32_t)labels[t];
        const int i = (int)(u >> 24), j = (int)((u >> 16) & 255u), k = (int)((u >> 8) & 255u),
                  l = (int)(u & 255u);
        int nq4;
        (void)sym_images(T, T.ishell[i - 1], T.ishell[j - 1], T.ishell[k - 1], T.ishell[l - 1], &nq4);
        double v = xx[t] * ((double)nq4 * dfnt);
        if (i == j) v = v + v;
        if (k == l) v = v + v;
        if (i == k && j == l) v = v + v;
        if (g > 0) {
            const int s = comp_sign(T, i - 1, g) * comp_sign(T, j - 1, g) * comp_sign(T, k - 1, g) *
                          comp_sign(T, l - 1, g);
            if (s < 0) v = -v;
        }
        out[p] = v;
    }
}

// ------------------------------------------------------------------ block sandwich kernel
// One half-transform of one symmetry block, as a batch of matrices Z_b (nr x ns):
//   T(r, c) = sum_s Z_b(r, s) CR(s, c)        (Q1 / Q3)
//   O(a, c) = sum_r CL(r, a) T(r, c)          (Q2 / Q4)
// with the output O(a, c) for a < ma, c < mc (a <= c when tri_out) at dst[row_b + pos],
// pos = tri_out ? pair0(c, a) : c * ma + a, pos < limit_b.
struct SymHalfDesc {
    const double* src;     // rows of the Z sources
    int64_t src_ld;        // doubles per source row
    const double* crR;     // CR padded in 16-column blocks of kCRows rows (cr_index)
    const double* crL;     // CL, the same layout
    const double* cqL;     // CL in the second-product order (cq_index)
    double* dst;
    int64_t dst_ld;        // half 1: doubles per output row
    int64_t nb;            // matrices
    int nr, ns, mc, ma;
    int ntiles;            // 64-column tiles of c per matrix
    int zrect;             // Z(r,s) = row[r * zld + s]; else row[pairsym(r + ro, s + so)]
    int zld, ro, so;
    int pmode;             // source row of matrix b: 0 = b; 1 = pairsym(offd + p, offc + q) for
                           // the AO pair b = pair0(p, q); 2 = the same with p = b / pn, q = b % pn
    int pn, offd, offc;
    int tri_out;
    int half;              // 1: row b at b * dst_ld, limit dst_ld; 2: as below
    int diag;              // half 2: triangular block, row b holds min(b + 1, nij) values;
                           // otherwise row b at b * nij, limit nij
    int64_t nij;
};

constexpr int kSymNW = 4;                      // waves per work-group (64 columns)
constexpr int kSymKC = 16;                     // Q1 panel depth

template <int NRB>
__device__ __forceinline__ void sym_tile(const SymHalfDesc& D, int64_t local) {
    constexpr int NT = kSymNW * 64;
    constexpr int EP = 16 * kSymKC;            // 256 elements per 16-row pass
    constexpr int NU = NRB;                    // one pass per thread per row block
    constexpr int RP = NRB * 16;
    constexpr int LDZ = RP + 1;
    constexpr int PANEL = kSymKC * LDZ;
    static_assert(NT == EP, "one staging sweep per 16-row pass");
    extern __shared__ double smem[];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int lm = lane & 15, lg = lane >> 4;

    const int64_t b = local / D.ntiles;
    const int tile = (int)(local - b * D.ntiles);
    int64_t srow;
    if (D.pmode == 0) {
        srow = b;
    } else {
        int64_t p, q;
        if (D.pmode == 1) {
            p = (int64_t)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
            while (pair0(p, 0) > b) --p;
            while (pair0(p + 1, 0) <= b) ++p;
            q = b - pair0(p, 0);
        } else {
            p = b / D.pn;
            q = b - p * D.pn;
        }
        srow = pairsym(D.offd + p, D.offc + q);
    }
    const double* __restrict__ zsrc = D.src + srow * D.src_ld;
    const int cw = 16 * (tile + D.ntiles * w);
    const bool wave_active = cw < D.mc;
    const int nr = D.nr, ns = D.ns;

    double zst[NU];
    auto zmap = [&](int u, int s0, int& r, int& kk) {
        const int e = tid;
        const bool above = !D.zrect && (16 * u + 15 + D.ro < s0 + D.so);
        r = 16 * u + (above ? (e & 15) : e / kSymKC);
        kk = above ? (e >> 4) : (e % kSymKC);
    };
    auto fetch = [&](int s0) {
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            int r, kk;
            zmap(u, s0, r, kk);
            const int s = s0 + kk;
            double v = 0.0;
            if (r < nr && s < ns) {
                const int64_t idx = D.zrect ? (int64_t)r * D.zld + s : pairsym(r + D.ro, s + D.so);
                v = zsrc[idx];
            }
            zst[u] = v;
        }
    };
    auto commit = [&](int s0, int boff) {
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            int r, kk;
            zmap(u, s0, r, kk);
            smem[boff + kk * LDZ + r] = zst[u];
        }
    };

    f64x4 acc[NRB];
#pragma unroll
    for (int i = 0; i < NRB; ++i) acc[i] = f64x4{0.0, 0.0, 0.0, 0.0};
    constexpr int NB = kSymKC / 4;
    const double* __restrict__ cbw = D.crR + (int64_t)(cw >> 4) * (kCRows * 16) + lg * 16 + lm;
    const int nchunk = (ns + kSymKC - 1) / kSymKC;
    fetch(0);
    commit(0, 0);
    __syncthreads();
    for (int ch = 0; ch < nchunk; ++ch) {
        const bool more = (ch + 1) < nchunk;
        if (more) fetch((ch + 1) * kSymKC);
        if (wave_active) {
            const int zo = (ch & 1) * PANEL;
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const double bv = cbw[(ch * kSymKC + 4 * k) * 16];
#pragma unroll
                for (int rb = 0; rb < NRB; ++rb) {
                    const double af = smem[zo + (4 * k + lg) * LDZ + rb * 16 + lm];
                    acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, bv, acc[rb], 0, 0, 0);
                }
            }
        }
        if (more) commit((ch + 1) * kSymKC, ((ch + 1) & 1) * PANEL);
        __syncthreads();
    }
    if (!wave_active) return;

    int64_t rowoff, limit;
    if (D.half == 1) {
        rowoff = b * D.dst_ld;
        limit = D.dst_ld;
    } else if (D.diag) {
        // row b holds min(b + 1, nij) values (tijkld.f:75-77 stops at nij = ilk)
        rowoff = b < D.nij ? pair0(b, 0) : pair0(D.nij, 0) + (b - D.nij) * D.nij;
        limit = b < D.nij ? b + 1 : D.nij;
    } else {
        rowoff = b * D.nij;
        limit = D.nij;
    }
    double* __restrict__ dst = D.dst + rowoff;
    const int last_ab = D.tri_out ? cw / 16 : (D.ma - 1) / 16;
    for (int ab = 0; ab <= last_ab; ++ab) {
        // a lane's four C values of row block rb: two 16-byte pieces of the cq_index copy
        typedef double f64x2 __attribute__((ext_vector_type(2)));
        const f64x2* __restrict__ cqa = (const f64x2*)(D.cqL + (int64_t)ab * 16 * 256) + lane;
        f64x4 o0 = f64x4{0.0, 0.0, 0.0, 0.0}, o1 = o0;
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb) {
            const f64x2 v0 = cqa[rb * 128], v1 = cqa[rb * 128 + 64];
            const double b0 = v0.x, b1 = v0.y, b2 = v1.x, b3 = v1.y;
            o0 = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[rb][0], b0, o0, 0, 0, 0);
            o1 = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[rb][1], b1, o1, 0, 0, 0);
            o0 = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[rb][2], b2, o0, 0, 0, 0);
            o1 = __builtin_amdgcn_mfma_f64_16x16x4f64(acc[rb][3], b3, o1, 0, 0, 0);
        }
        const int a = ab * 16 + lm;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = cw + lg + 4 * q;
            if (c < D.mc && a < D.ma && (!D.tri_out || a <= c)) {
                const int64_t pos = D.tri_out ? pair0(c, a) : (int64_t)c * D.ma + a;
                if (pos < limit) dst[pos] = o0[q] + o1[q];
            }
        }
    }
}

template <int NRB>
__global__ __launch_bounds__(kSymNW * 64) void k_sym_half(const SymHalfDesc* __restrict__ descs,
                                                          const int64_t* __restrict__ item_start,
                                                          int nblk) {
    const int64_t total = item_start[nblk];
    const int64_t it = xcd_item(blockIdx.x);
    if (it >= total) return;
    int lo = 0, hi = nblk;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (item_start[mid] <= it) lo = mid; else hi = mid;
    }
    const SymHalfDesc D = descs[lo];
    sym_tile<NRB>(D, it - item_start[lo]);
}

template <int NRB>
static hipError_t launch_sym_half_rows(const SymHalfDesc* d, const int64_t* starts, int nblk,
                                       int64_t items, hipStream_t st) {
    const size_t shmem = sizeof(double) * 2 * kSymKC * (NRB * 16 + 1);
    static bool attr_set = false;                  // 16 row blocks need 65.8 KB of LDS
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)k_sym_half<NRB>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    const int64_t grid = xcd_grid(items);
    if (grid <= 0 || grid > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sym_half<NRB>, dim3((unsigned)grid), dim3(kSymNW * 64), shmem, st, d,
                       starts, nblk);
    return hipGetLastError();
}

static hipError_t launch_sym_half(int nrb, const SymHalfDesc* d, const int64_t* starts, int nblk,
                                  int64_t items, hipStream_t st) {
    switch (nrb) {
#define SYM_CASE(R) case R: return launch_sym_half_rows<R>(d, starts, nblk, items, st);
        SYM_CASE(1) SYM_CASE(2) SYM_CASE(3) SYM_CASE(4) SYM_CASE(5) SYM_CASE(6) SYM_CASE(7)
        SYM_CASE(8) SYM_CASE(9) SYM_CASE(10) SYM_CASE(11) SYM_CASE(12) SYM_CASE(13) SYM_CASE(14)
        SYM_CASE(15) SYM_CASE(16)
#undef SYM_CASE
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cipsi

using namespace cipsi;

// ------------------------------------------------------------------ host orchestration
namespace {

struct HostBlock {
    int ii, jj, kk, ll;            // irreps (1-based)
    int nij, nkl, ijkls;
    bool yij, ykl, yijkl, ypq, yrs;  // MO pairs of (A,B) / (C,D) triangular, diagonal block,
                                     // AO pairs of (A,B) / (C,D) triangular (same class)
    int64_t npq, nlk, nijc;          // AO pairs of (C,D); MO pairs of (A,B); of (C,D)
    int64_t h_off, out_off, out_cnt;
};

int fail(cipsi_ijkl_ctx* ctx, int rc, const std::string& why) {
    if (ctx) ctx->last_error = why;
    return rc;
}

inline int64_t npr(int64_t n) { return n * (n + 1) / 2; }
inline int64_t even(int64_t x) { return (x + 31) & ~int64_t(31); }

// itijkl.f:28-36 (its table ltabs): the type of a block reached through a transformation
// whose index permutation is perm (2 + bit 0: i <-> j, bit 1: k <-> l, bit 2: the pairs
// exchanged) from a block of type kt.  Types 2..9 compose as the permutations they stand
// for; 1 and 10 (blocks stored as triangles) absorb them, 1 with 10 giving 10.
inline int block_type(int perm, int kt) {
    if (perm == 1 || perm == 10 || kt == 1 || kt == 10)
        return (perm == 1 && kt == 10) || (perm == 10 && kt == 1) ? 10 : 1;
    return (((perm - 2) ^ (kt - 2)) & 7) + 2;
}

}  // namespace

extern "C" int cipsi_ijkl_sym_transform_labels(
    cipsi_ijkl_ctx* ctx, const cipsi_ijkl_sym* sym, const double* xx, const int32_t* labels,
    int64_t n_entries, double* mo_stream, int64_t mo_capacity, int64_t* mo_count, int* ipair,
    int max_pairs, int* npaire, int16_t* nad, int16_t* ktypsy) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if (!sym || !mo_count || !npaire || n_entries < 0 || (n_entries > 0 && (!xx || !labels)))
        return CIPSI_ERR_ARG;
    const int n = sym->n_ao, nsym = sym->nsym, nt = sym->nt, nshell = sym->nshell;
    if (n < 1 || nsym < 1 || nt < 1 || nshell < 1 || !sym->its || !sym->isymp || !sym->nbp ||
        !sym->nbo || !sym->ind || !sym->ivecsy || !sym->c || !sym->ishell || !sym->kloc ||
        !sym->ktyp || !sym->numc || !sym->newsh || !sym->iptr || !sym->idtr || !sym->iftr)
        return CIPSI_ERR_ARG;
    if (n > kSymMaxAO) return fail(ctx, CIPSI_ERR_LABEL, "sym: N_AO > 255 (8-bit label fields)");
    if (nsym > kSymMaxIrreps || nt > kSymMaxOps || nshell > kSymMaxShell)
        return fail(ctx, CIPSI_ERR_UNSUPPORTED, "sym: more than 14 irreps or 8 operations");
    const int ntrsy = sym->ntrsy;
    if (ntrsy < 0 || ntrsy > kSymMaxTrans || (ntrsy > 0 && !sym->isydeg))
        return fail(ctx, CIPSI_ERR_ARG, "sym: bad ntrsy / isydeg");
    for (int k = 0; k < nsym * ntrsy; ++k)
        if (sym->isydeg[k] < -nsym || sym->isydeg[k] > nsym)
            return fail(ctx, CIPSI_ERR_ARG, "sym: an isydeg entry is not a signed irrep");
    if (n_entries >= (int64_t)1 << 60) return CIPSI_ERR_ARG;
    // ---- classes, AO subsets (vector.f numbering), the class-ordered AO basis
    const int* isymp = sym->isymp;
    const int nclass = isymp[nsym - 1];
    for (int s = 0; s < nsym; ++s) {
        if (isymp[s] < 1 || isymp[s] > nclass || (s > 0 && isymp[s] < isymp[s - 1]))
            return fail(ctx, CIPSI_ERR_ARG, "sym: isymp must be non-decreasing classes 1..nsymp");
        if (sym->nbp[s] < 0 || sym->nbo[s] < 0 || sym->nbp[s] > n)
            return fail(ctx, CIPSI_ERR_ARG, "sym: bad nbp / nbo");
        for (int t = 0; t < nsym; ++t) {
            const int v = sym->its[s + t * nsym];
            if (v < 1 || v > nsym) return fail(ctx, CIPSI_ERR_ARG, "sym: bad product table its");
        }
    }
    SymTables T;
    std::memset(&T, 0, sizeof(T));
    T.n = n; T.nt = nt; T.nshell = nshell; T.nclass = nclass;
    std::vector<int> cls_of(n, -1), cls_first(nclass, -1), cls_size(nclass, 0);
    for (int s = 0; s < nsym; ++s) {
        const int c = isymp[s] - 1;
        int k = 0;
        for (int a = 0; a < n; ++a) {
            const int v = sym->ivecsy[s * n + a];
            if (v == 0) continue;
            if (v != ++k) return fail(ctx, CIPSI_ERR_ARG, "sym: ivecsy is not vector.f's numbering");
            if (cls_of[a] >= 0 && cls_of[a] != c)
                return fail(ctx, CIPSI_ERR_UNSUPPORTED, "sym: an AO belongs to two classes");
            cls_of[a] = c;
        }
        if (k != sym->nbp[s]) return fail(ctx, CIPSI_ERR_ARG, "sym: nbp does not match ivecsy");
        if (cls_first[c] < 0) {
            cls_first[c] = s;
            cls_size[c] = k;
        } else {
            for (int a = 0; a < n; ++a)
                if ((sym->ivecsy[s * n + a] != 0) != (sym->ivecsy[cls_first[c] * n + a] != 0))
                    return fail(ctx, CIPSI_ERR_ARG, "sym: irreps of one class differ in AOs");
        }
    }
    std::vector<int> cls_off(nclass + 1, 0);
    for (int c = 0; c < nclass; ++c) cls_off[c + 1] = cls_off[c] + cls_size[c];
    {
        std::vector<int> fill(nclass, 0);
        int extra = cls_off[nclass];
        for (int a = 0; a < n; ++a) {
            T.cls[a] = cls_of[a];
            T.perm[a] = cls_of[a] >= 0 ? cls_off[cls_of[a]] + fill[cls_of[a]]++ : extra++;
        }
    }
    for (int s = 0; s < nshell; ++s) {
        const int k0 = sym->kloc[s] - 1;
        if (k0 < 0 || k0 >= n) return fail(ctx, CIPSI_ERR_ARG, "sym: bad kloc");
        T.kloc[s] = k0;
        for (int g = 0; g < nt; ++g) {
            const int im = sym->newsh[s + g * nshell] - 1;
            if (im < 0 || im >= nshell) return fail(ctx, CIPSI_ERR_ARG, "sym: bad newsh");
            T.newsh[s * kSymMaxOps + g] = im;
        }
    }
    for (int a = 0; a < n; ++a) {
        const int sh = sym->ishell[a] - 1, kt = sym->ktyp[a], c = sym->numc[a];
        if (sh < 0 || sh >= nshell || kt < 1 || kt > 4 || c < 0 || c > 9 ||
            (kt == 2 && c > 2) || (kt == 3 && c > 5))
            return fail(ctx, CIPSI_ERR_ARG, "sym: bad ishell / ktyp / numc");
        T.ishell[a] = sh; T.ktyp[a] = kt; T.numc[a] = c;
        for (int g = 0; g < nt; ++g)
            if (T.kloc[T.newsh[sh * kSymMaxOps + g]] + c >= n)
                return fail(ctx, CIPSI_ERR_ARG, "sym: an image AO lies outside the basis");
    }
    for (int g = 0; g < nt; ++g) {
        for (int c = 0; c < 3; ++c) T.sgn_p[c * kSymMaxOps + g] = sym->iptr[c + 3 * g] < 0 ? -1 : 1;
        for (int c = 0; c < 6; ++c) T.sgn_d[c * kSymMaxOps + g] = sym->idtr[c + 6 * g] < 0 ? -1 : 1;
        for (int c = 0; c < 10; ++c) T.sgn_f[c * kSymMaxOps + g] = sym->iftr[c + 10 * g] < 0 ? -1 : 1;
    }
    const int64_t pa = npr(n), nblkpos = npr(pa);
    hipStream_t st = ctx->stream;

    // ---- device buffers: tables + flags, labels, values, owners, packed block, pair matrix
    const size_t small_bytes = 4096 + sizeof(SymTables) + sizeof(int) * kSymMaxClassKeys;
    if ((rc = ctx_alloc(ctx, ctx->sym_small, small_bytes))) return rc;
    unsigned long long* d_sc = (unsigned long long*)ctx->sym_small.ptr;
    SymTables* d_T = (SymTables*)((char*)ctx->sym_small.ptr + 256);
    int* d_touched = (int*)((char*)ctx->sym_small.ptr + 256 + ((sizeof(SymTables) + 255) & ~size_t(255)));
    const int64_t ne = n_entries > 0 ? n_entries : 1;
    if ((rc = ctx_alloc(ctx, ctx->io_in, sizeof(double) * ne))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->io_in2, sizeof(int32_t) * ne))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->sym_owner, sizeof(unsigned long long) * nblkpos))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->packed, sizeof(double) * nblkpos))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->vsq, sizeof(double) * pa * pa))) return rc;
    const double* d_xx = (const double*)ctx->io_in.ptr;
    const int32_t* d_lab = (const int32_t*)ctx->io_in2.ptr;
    unsigned long long* d_owner = (unsigned long long*)ctx->sym_owner.ptr;
    double* d_v = (double*)ctx->packed.ptr;
    double* d_vsq = (double*)ctx->vsq.ptr;
    if (n_entries > 0) {
        if ((rc = ctx_copy_h2d(ctx, ctx->io_in.ptr, xx, sizeof(double) * n_entries))) return rc;
        if ((rc = ctx_copy_h2d(ctx, ctx->io_in2.ptr, labels, sizeof(int32_t) * n_entries))) return rc;
    }
#define SYM_HIP(expr)                                                      \
    do {                                                                   \
        hipError_t _e = (expr);                                            \
        if (_e != hipSuccess) {                                            \
            (void)hipGetLastError();                                       \
            return fail(ctx, _e == hipErrorOutOfMemory ? CIPSI_ERR_NOMEM : CIPSI_ERR_HIP, \
                        std::string(#expr) + ": " + hipGetErrorString(_e)); \
        }                                                                  \
    } while (0)
    SYM_HIP(hipMemcpyAsync(d_T, &T, sizeof(T), hipMemcpyHostToDevice, st));
    SYM_HIP(hipMemsetAsync(d_touched, 0, sizeof(int) * kSymMaxClassKeys, st));
    SYM_HIP(hipMemsetAsync(d_owner, 0, sizeof(unsigned long long) * nblkpos, st));
    {
        StageTimer tm(ctx, "sym_unpack");
        hipLaunchKernelGGL(k_sym_scan_init, dim3(1), dim3(64), 0, st, d_sc, n_entries);
        if (n_entries > 0) {
            const int64_t g = std::min<int64_t>((n_entries + 255) / 256, 8192);
            hipLaunchKernelGGL(k_sym_scan, dim3((unsigned)g), dim3(256), 0, st, d_lab, n_entries, n, d_sc);
            hipLaunchKernelGGL(k_sym_claim, dim3((unsigned)g), dim3(256), 0, st, d_lab, d_T, d_sc,
                               d_owner, d_touched);
        }
        const int64_t gr = std::min<int64_t>((nblkpos + 255) / 256, 65536);
        hipLaunchKernelGGL(k_sym_resolve, dim3((unsigned)gr), dim3(256), 0, st, d_xx, d_lab, d_T,
                           d_owner, nblkpos, d_v);
        SYM_HIP(hipGetLastError());
    }
    unsigned long long hsc[2] = {0ull, 0ull};
    std::vector<int> touched(kSymMaxClassKeys, 0);
    SYM_HIP(hipMemcpyAsync(hsc, d_sc, sizeof(hsc), hipMemcpyDeviceToHost, st));
    SYM_HIP(hipMemcpyAsync(touched.data(), d_touched, sizeof(int) * kSymMaxClassKeys,
                           hipMemcpyDeviceToHost, st));
    SYM_HIP(hipStreamSynchronize(st));
    if (hsc[1] < hsc[0]) return fail(ctx, CIPSI_ERR_LABEL, "sym: a label field lies outside 1..N_AO");

    // ---- the block list in itijkl's order (itijkl.f:70-247), with the directory
    int nsyt = nsym * (nsym + 1) / 2;
    nsyt = std::min(2000, nsyt * (nsyt + 1) / 2);
    std::vector<int> h_nad(nsyt, 0), h_kt(nsyt, 1);
    std::vector<HostBlock> blocks;
    auto its = [&](int a, int b) { return sym->its[(a - 1) + (b - 1) * nsym]; };
    const int* nbp = sym->nbp;
    const int* nbo = sym->nbo;
    for (int msi = 1; msi <= nclass; ++msi)
        for (int msj = 1; msj <= nclass; ++msj)
            for (int msk = 1; msk <= nclass; ++msk)
                for (int msl = 1; msl <= nclass; ++msl) {
                    const bool ypq = msi == msj, yrs = msk == msl;
                    bool yr = true, ysaut = false;
                    for (int ii = 1; ii <= nsym; ++ii)
                        for (int jj = 1; jj <= ii; ++jj) {
                            int64_t mono1 = (int64_t)nbp[ii - 1] * nbp[jj - 1];
                            if (ypq) mono1 = (mono1 + nbp[jj - 1]) / 2;
                            const bool yij = ii == jj;
                            const int nij = ii * (ii - 1) / 2 + jj;
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
                                    const bool ykl = kk == ll;
                                    bool yijkl = ii == kk && jj == ll;
                                    if (its(its(kk, ll), its(ii, jj)) != 1) continue;
                                    const int nkl = kk * (kk - 1) / 2 + ll;
                                    int64_t mkl = (int64_t)nbo[kk - 1] * nbo[ll - 1];
                                    if (ykl) mkl = (mkl + nbo[kk - 1]) / 2;
                                    if (mkl * mij <= 0) continue;
                                    const int ijkls = nij * (nij - 1) / 2 + nkl;
                                    // itijkl.f:121-200: the transformations' images of the
                                    // quartet, canonically ordered; one that comes earlier
                                    // makes this block redundant (entry -> that block's),
                                    // one that comes later and is forbidden, forbidden
                                    bool skip = false;
                                    for (int t = 0; t < ntrsy && !skip; ++t) {
                                        const int* map = sym->isydeg + (size_t)t * nsym;
                                        int q[4] = {map[ii - 1], map[jj - 1], map[kk - 1], map[ll - 1]};
                                        if (!q[0] || !q[1] || !q[2] || !q[3]) continue;
                                        int sgn = 1;
                                        for (int& x : q)
                                            if (x < 0) { x = -x; sgn = -sgn; }
                                        if (q[0] == kk && q[1] == ll && q[2] == ii && q[3] == jj) yijkl = true;
                                        int perm = 2;
                                        if (q[0] < q[1]) { std::swap(q[0], q[1]); perm += 1; }
                                        if (q[2] < q[3]) { std::swap(q[2], q[3]); perm += 2; }
                                        if (q[0] < q[2] || (q[0] == q[2] && q[3] > q[1])) {
                                            std::swap(q[0], q[2]);
                                            std::swap(q[1], q[3]);
                                            perm += 4;
                                        }
                                        const int ijt = q[0] * (q[0] - 1) / 2 + q[1];
                                        const int klt = q[2] * (q[2] - 1) / 2 + q[3];
                                        if (ijt > nij || (ijt == nij && klt >= nkl)) {
                                            skip = its(its(q[0], q[1]), its(q[2], q[3])) != 1;
                                            continue;
                                        }
                                        if (ijkls > nsyt)
                                            return fail(ctx, CIPSI_ERR_UNSUPPORTED, "sym: directory beyond nsyt entries");
                                        const int ijklt = ijt * (ijt - 1) / 2 + klt;
                                        const int kt_t = h_kt[ijklt - 1];
                                        if (kt_t == 1 || kt_t == 10) perm = perm == 5 ? 10 : 1;
                                        h_kt[ijkls - 1] = block_type(perm, kt_t);
                                        h_nad[ijkls - 1] = sgn * h_nad[ijklt - 1];
                                        skip = true;
                                    }
                                    if (skip) continue;
                                    if (yr) {
                                        // transd fills the class quartet's block; it finds no
                                        // integral of this class quartet -> skipped
                                        yr = false;
                                        const int p1 = (msi - 1) * msi / 2 + (msj - 1);
                                        const int p2 = (msk - 1) * msk / 2 + (msl - 1);
                                        const int key = p1 >= p2 ? p1 * (p1 + 1) / 2 + p2 : p2 * (p2 + 1) / 2 + p1;
                                        ysaut = touched[key] == 0;
                                    }
                                    if (ysaut) continue;
                                    if (ijkls > nsyt)
                                        return fail(ctx, CIPSI_ERR_UNSUPPORTED, "sym: directory beyond nsyt entries");
                                    HostBlock B{};
                                    B.ii = ii; B.jj = jj; B.kk = kk; B.ll = ll;
                                    B.nij = nij; B.nkl = nkl; B.ijkls = ijkls;
                                    B.yij = yij; B.ykl = ykl; B.yijkl = yijkl; B.ypq = ypq; B.yrs = yrs;
                                    B.npq = mono2;
                                    B.nlk = mij;
                                    B.nijc = mkl;
                                    blocks.push_back(B);
                                    h_kt[ijkls - 1] = yijkl ? 1 : 2;
                                    h_nad[ijkls - 1] = (int)blocks.size();
                                }
                            }
                        }
                }
    const int nb = (int)blocks.size();
    if (ipair && max_pairs < nb) return fail(ctx, CIPSI_ERR_ARG, "sym: max_pairs too small");
    int64_t hsz = 0, total = 0;
    for (auto& B : blocks) {
        B.h_off = hsz;
        hsz += even(B.npq * B.nlk);
        B.out_off = total;
        // a triangular block: row ilk holds min(ilk, nijc) values (tijkld.f:75-77)
        B.out_cnt = !B.yijkl ? B.nlk * B.nijc
                  : B.nijc >= B.nlk ? npr(B.nlk) : npr(B.nijc) + (B.nlk - B.nijc) * B.nijc;
        total += B.out_cnt;
    }
    *mo_count = total;
    *npaire = nb;
    if (ipair)
        for (int b = 0; b < nb; ++b) { ipair[2 * b] = blocks[b].nij; ipair[2 * b + 1] = blocks[b].nkl; }
    for (int i = 0; i < nsyt; ++i) {
        if (nad) nad[i] = (int16_t)h_nad[i];
        if (ktypsy) ktypsy[i] = (int16_t)h_kt[i];
    }
    if (!mo_stream) return CIPSI_OK;                 // size query
    if (mo_capacity < total) return fail(ctx, CIPSI_ERR_ARG, "sym: mo_capacity below the stream length");
    if (nb == 0) return CIPSI_OK;

    // ---- C blocks (vector.f's /om/ c) padded per irrep
    int64_t csz = 0;
    std::vector<int64_t> c_off(nsym, 0), cr_off(nsym, 0);
    int64_t crsz = 0;
    for (int s = 0; s < nsym; ++s) {
        c_off[s] = sym->ind[s] - 1;
        csz = std::max<int64_t>(csz, c_off[s] + (int64_t)nbp[s] * nbo[s]);
        cr_off[s] = crsz;
        crsz += cr_pair_size((nbo[s] + 15) & ~15) + 32;
    }
    if ((rc = ctx_alloc(ctx, ctx->io_c, sizeof(double) * std::max<int64_t>(csz, 1)))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->cr, sizeof(double) * crsz))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->h, sizeof(double) * std::max<int64_t>(hsz, 1)))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->sym_ht, sizeof(double) * std::max<int64_t>(hsz, 1)))) return rc;
    if ((rc = ctx_alloc(ctx, ctx->io_out, sizeof(double) * total))) return rc;
    SYM_HIP(hipMemcpyAsync(ctx->io_c.ptr, sym->c, sizeof(double) * csz, hipMemcpyHostToDevice, st));
    double* d_cr = (double*)ctx->cr.ptr;
    for (int s = 0; s < nsym; ++s) {
        if (nbo[s] == 0 || nbp[s] == 0) continue;
        SYM_HIP(launch_pad_c((const double*)ctx->io_c.ptr + c_off[s], nbp[s], nbp[s], nbo[s], kCRows,
                             (nbo[s] + 15) & ~15, d_cr + cr_off[s], st));
    }
    {
        StageTimer tm(ctx, "expand");
        SYM_HIP(launch_expand_packed(d_v, pa, d_vsq, st));
    }

    // ---- descriptors of both halves, grouped by row class; transposes
    double* d_h = (double*)ctx->h.ptr;
    double* d_ht = (double*)ctx->sym_ht.ptr;
    double* d_out = (double*)ctx->io_out.ptr;
    std::vector<SymHalfDesc> d1(nb), d2(nb);
    std::vector<LayoutGeom> tg(nb);
    int64_t ttiles = 0;
    for (int b = 0; b < nb; ++b) {
        const HostBlock& B = blocks[b];
        const int A = B.ii - 1, Bi = B.jj - 1, C = B.kk - 1, D = B.ll - 1;
        SymHalfDesc& x = d1[b];
        x.src = d_vsq; x.src_ld = pa;
        x.crR = d_cr + cr_off[A]; x.crL = d_cr + cr_off[Bi];
        x.cqL = x.crL + (int64_t)kCRows * ((nbo[Bi] + 15) & ~15);
        x.dst = d_h + B.h_off; x.dst_ld = B.nlk;
        x.nb = B.npq;
        x.nr = nbp[Bi]; x.ns = nbp[A]; x.mc = nbo[A]; x.ma = nbo[Bi];
        x.ntiles = (x.mc + 63) / 64;
        x.zrect = 0; x.zld = 0;
        x.ro = cls_off[isymp[Bi] - 1]; x.so = cls_off[isymp[A] - 1];
        x.pmode = B.yrs ? 1 : 2; x.pn = nbp[C];
        x.offd = cls_off[isymp[D] - 1]; x.offc = cls_off[isymp[C] - 1];
        x.tri_out = B.yij; x.half = 1; x.diag = 0; x.nij = 0;
        SymHalfDesc& y = d2[b];
        y.src = d_ht + B.h_off; y.src_ld = B.npq;
        y.crR = d_cr + cr_off[C]; y.crL = d_cr + cr_off[D];
        y.cqL = y.crL + (int64_t)kCRows * ((nbo[D] + 15) & ~15);
        y.dst = d_out + B.out_off; y.dst_ld = 0;
        y.nb = B.nlk;
        y.nr = nbp[D]; y.ns = nbp[C]; y.mc = nbo[C]; y.ma = nbo[D];
        y.ntiles = (y.mc + 63) / 64;
        y.zrect = B.yrs ? 0 : 1; y.zld = nbp[C]; y.ro = 0; y.so = 0;
        y.pmode = 0; y.pn = 0; y.offd = 0; y.offc = 0;
        y.tri_out = B.ykl; y.half = 2; y.diag = B.yijkl; y.nij = B.nijc;
        LayoutGeom& g = tg[b];
        std::memset(&g, 0, sizeof(g));
        g.h = d_h + B.h_off; g.vsq = d_ht + B.h_off; g.pa = B.npq; g.pm = B.nlk; g.ldv = B.npq;
        g.trans_start = ttiles;
        ttiles += transpose_tiles(B.npq, B.nlk);
    }
    // order the descriptors of each half by row class; item tables per class
    auto rows_class = [](int nr) { return std::max(1, (nr + 15) / 16); };
    struct Group { int nrb; int first, count; int64_t items; };
    auto build = [&](std::vector<SymHalfDesc>& d, std::vector<int64_t>& starts, std::vector<Group>& groups) {
        std::vector<int> idx(nb);
        for (int b = 0; b < nb; ++b) idx[b] = b;
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) { return rows_class(d[a].nr) < rows_class(d[c].nr); });
        std::vector<SymHalfDesc> sorted(nb);
        for (int b = 0; b < nb; ++b) sorted[b] = d[idx[b]];
        d.swap(sorted);
        starts.clear();
        groups.clear();
        for (int b = 0; b < nb;) {
            const int r = rows_class(d[b].nr);
            Group G{r, b, 0, 0};
            while (b < nb && rows_class(d[b].nr) == r) {
                starts.push_back(G.items);
                G.items += d[b].nb * d[b].ntiles;
                ++G.count;
                ++b;
            }
            starts.push_back(G.items);
            groups.push_back(G);
        }
    };
    std::vector<int64_t> s1, s2;
    std::vector<Group> g1, g2;
    build(d1, s1, g1);
    build(d2, s2, g2);
    if (std::max(d1.size(), d2.size()) > 0 && (g1.back().nrb > 16 || g2.back().nrb > 16))
        return fail(ctx, CIPSI_ERR_UNSUPPORTED, "sym: an irrep with more than 256 AOs");
    const size_t tab_bytes = sizeof(SymHalfDesc) * 2 * nb + sizeof(int64_t) * (s1.size() + s2.size()) +
                             sizeof(LayoutGeom) * nb + 1024;
    if ((rc = ctx_alloc(ctx, ctx->sym_tab, tab_bytes))) return rc;
    char* tb = (char*)ctx->sym_tab.ptr;
    SymHalfDesc* dd1 = (SymHalfDesc*)tb;
    SymHalfDesc* dd2 = dd1 + nb;
    LayoutGeom* dtg = (LayoutGeom*)(((uintptr_t)(dd2 + nb) + 255) & ~uintptr_t(255));
    int64_t* ds1 = (int64_t*)(((uintptr_t)(dtg + nb) + 255) & ~uintptr_t(255));
    int64_t* ds2 = ds1 + s1.size();
    SYM_HIP(hipMemcpyAsync(dd1, d1.data(), sizeof(SymHalfDesc) * nb, hipMemcpyHostToDevice, st));
    SYM_HIP(hipMemcpyAsync(dd2, d2.data(), sizeof(SymHalfDesc) * nb, hipMemcpyHostToDevice, st));
    SYM_HIP(hipMemcpyAsync(dtg, tg.data(), sizeof(LayoutGeom) * nb, hipMemcpyHostToDevice, st));
    SYM_HIP(hipMemcpyAsync(ds1, s1.data(), sizeof(int64_t) * s1.size(), hipMemcpyHostToDevice, st));
    SYM_HIP(hipMemcpyAsync(ds2, s2.data(), sizeof(int64_t) * s2.size(), hipMemcpyHostToDevice, st));
    auto run_groups = [&](SymHalfDesc* dd, int64_t* ds, const std::vector<Group>& groups) -> hipError_t {
        int64_t soff = 0;
        for (const Group& G : groups) {
            if (G.items > 0) {
                hipError_t e = launch_sym_half(G.nrb, dd + G.first, ds + soff, G.count, G.items, st);
                if (e != hipSuccess) return e;
            }
            soff += G.count + 1;
        }
        return hipSuccess;
    };
    {
        StageTimer tm(ctx, "half1");
        SYM_HIP(run_groups(dd1, ds1, g1));
    }
    {
        StageTimer tm(ctx, "transpose");
        if (ttiles > 0) SYM_HIP(launch_group_transpose(dtg, nb, ttiles, st));
    }
    {
        StageTimer tm(ctx, "half2");
        SYM_HIP(run_groups(dd2, ds2, g2));
    }
    SYM_HIP(hipMemcpyAsync(mo_stream, d_out, sizeof(double) * total, hipMemcpyDeviceToHost, st));
    SYM_HIP(hipStreamSynchronize(st));
#undef SYM_HIP
    return CIPSI_OK;
}
