// Host-side preparation of the MO coefficients of a molecule with point-group symmetry,
// the steps the reference ijkl takes before itijkl (ijkl.f:186-350):
//
//   group tables   otodon.f:163-317 with the tables of BLOCKDATA1.f: irreps, their
//                  product table and characters for the non-degenerate groups
//   symmetry AOs   otodon.f:320-353: per irrep, which AOs carry its symmetry orbitals
//   linear groups  otodon.f:36-144: C-inf-v and D-inf-h, whose AOs ijkl classifies by
//                  their angular character about the axis (and g / u for D-inf-h), with
//                  the product table and the two symmetry transformations (rotations by
//                  90 and 45 degrees) that itijkl uses to skip equivalent blocks
//   rotat2         rotat2.f:50-423: frozen-column compaction; the symmetry vectors
//                  normalised with S; classes of irreps with the same AOs, made contiguous;
//                  levels of degenerate MOs (the MO diagonal of h, tmono.f:37-54); the
//                  irrep of every level by overlap with the symmetry vectors, projecting
//                  degenerate levels onto their irreps; normalisation and orthogonality
//                  checks (error tags 2, 1, -1); MOs reordered occupied-by-irrep then
//                  virtual-by-irrep
//   vector         vector.f:17-46: per-irrep packing of C on the irrep's AOs
//
// Host FP64 code built without fused multiply-add, summing in the reference's order, so
// the coefficients equal the reference's bit for bit (its ijom file).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/cipsi_ijkl.h"

namespace {

// Characters chi(irrep, operation) of the groups otodon handles without degeneracies,
// operations in HONDO's order (the jsym2 / jsym4 / jsym8 tables of BLOCKDATA1.f).
const int kChi2[2][2] = {{1, 1}, {1, -1}};
const int kChi4[4][4] = {{1, 1, 1, 1}, {1, 1, -1, -1}, {1, -1, 1, -1}, {1, -1, -1, 1}};
const int kChi8[8][8] = {
    {1, 1, 1, 1, 1, 1, 1, 1},     {1, 1, -1, -1, 1, 1, -1, -1}, {1, -1, 1, -1, 1, -1, 1, -1},
    {1, -1, -1, 1, 1, -1, -1, 1}, {1, 1, 1, 1, -1, -1, -1, -1}, {1, 1, -1, -1, -1, -1, 1, 1},
    {1, -1, 1, -1, -1, 1, -1, 1}, {1, -1, -1, 1, -1, 1, 1, -1}};

inline int64_t tri(int64_t i) { return i * (i - 1) / 2; }      // 1-based pair offset

struct Packed {                                                 // packed symmetric, 1-based
    const double* p;
    double operator()(int i, int j) const { return p[(i >= j ? tri(i) + j : tri(j) + i) - 1]; }
};

bool same_name(const char* g, const char* name) {
    int k = 0;
    for (; name[k]; ++k) {
        char c = g[k];
        if (c >= 'a' && c <= 'z') c = (char)(c - 'a' + 'A');
        if (c != name[k]) return false;
    }
    for (; k < 8 && g[k]; ++k)
        if (g[k] != ' ') return false;
    return true;
}

}  // namespace

extern "C" int cipsi_ijkl_sym_group(const char* group, int naxis, int nt, int* nsym, int* its,
                                    int* jsym) {
    if (!group || !nsym || !its || !jsym) return CIPSI_ERR_ARG;
    int ns = 0;
    if (same_name(group, "CINFV") || same_name(group, "DINFH")) return CIPSI_ERR_UNSUPPORTED;
    if (same_name(group, "C1")) ns = 1;
    else {
        const bool cs_ci = same_name(group, "CS") || same_name(group, "CI");
        if (naxis != 2 && !cs_ci) return CIPSI_ERR_UNSUPPORTED;       // otodon.f:276-280
        if (same_name(group, "CN") || cs_ci) ns = 2;
        else if (same_name(group, "DN") || same_name(group, "CNV") || same_name(group, "CNH")) ns = 4;
        else if (same_name(group, "DNH")) ns = 8;
        else return CIPSI_ERR_ARG;                                      // otodon.f:314-316
    }
    if (nt != ns) return CIPSI_ERR_ARG;             // one operation per irrep in these groups
    *nsym = ns;
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j < ns; ++j) {
            its[i + j * ns] = (i ^ j) + 1;          // the product tables its2 / its4 / its8
            const int c = ns == 1 ? 1 : ns == 2 ? kChi2[i][j] : ns == 4 ? kChi4[i][j] : kChi8[i][j];
            jsym[i + j * ns] = c;
        }
    return CIPSI_OK;
}

namespace {

// rotat2.f:50-423 and vector.f:17-46 for the symmetry-orbital table vs (nsym x n, the
// signed orbital number of each AO per irrep, as otodon leaves it in ivecsy).
int rotat2_vector(int n, int nsym, std::vector<int>& vs, const double* s_packed,
                  const double* h_packed, const double* c_square, int nvecf, const int* ygel,
                  int noc, double trec, double tdeg, double* c_out, int* itsyv, int* isymp,
                  int* nsymp, int* ivecsy, int* nbp, int* nbo, int* ind, double* c_packed,
                  int* tag) {
    const int64_t nn = (int64_t)n * n;
    // ---- rotat2.f:50-60: active columns to the front (later columns stay as they were)
    std::vector<double> c(c_square, c_square + nn);
    {
        int act = 0;
        for (int i = 0; i < nvecf; ++i) {
            if (ygel && ygel[i]) continue;
            if (act != i) std::memcpy(&c[(size_t)act * n], &c[(size_t)i * n], sizeof(double) * n);
            ++act;
        }
    }
    std::vector<double> unit;
    if (!s_packed) {                                   // yatom: S = 1 (rotat2.f:79-85)
        unit.assign((size_t)n * (n + 1) / 2, 0.0);
        for (int i = 1; i <= n; ++i) unit[(size_t)tri(i) + i - 1] = 1.0;
        s_packed = unit.data();
    }
    const Packed S{s_packed}, H{h_packed};
    // ---- rotat2.f:91-112: symmetry vectors normalised with S; AO membership
    std::vector<double> vec((size_t)nsym * n);
    for (size_t k = 0; k < vec.size(); ++k) vec[k] = (double)vs[k];
    for (int ir = 0; ir < nsym; ++ir) {
        double* v = &vec[(size_t)ir * n];
        double norm = 0;
        for (int i = 1; i <= n; ++i)
            for (int j = 1; j <= n; ++j) norm = norm + v[i - 1] * S(i, j) * v[j - 1];
        if (norm < 1.e-8) norm = 1.0;
        norm = 1.0 / std::sqrt(norm);
        for (int i = 0; i < n; ++i) v[i] = v[i] * norm;
        for (int i = 0; i < n; ++i) vs[(size_t)ir * n + i] = std::fabs(v[i]) < 1.e-8 ? 0 : ir + 1;
    }
    // ---- rotat2.f:114-130: classes of irreps with the same AOs
    int ncls = 1;
    isymp[0] = 1;
    for (int ir = 1; ir < nsym; ++ir) {
        int found = -1;
        for (int jr = 0; jr < ir && found < 0; ++jr) {
            bool same = true;
            for (int a = 0; a < n && same; ++a)
                same = (vs[(size_t)jr * n + a] == 0) == (vs[(size_t)ir * n + a] == 0);
            if (same) found = jr;
        }
        isymp[ir] = found >= 0 ? isymp[found] : ++ncls;
    }
    // ---- rotat2.f:131-157: irreps moved so that each class is contiguous (its unchanged)
    std::vector<int> mo_irrep((size_t)nvecf, 0);       // itsyv of the namelist: 0
    {
        int pos = 0;
        for (int cl = 1; cl <= ncls; ++cl)
            for (int ir = cl - 1; ir < nsym; ++ir) {
                if (isymp[ir] != cl) continue;
                const int dst = pos++;
                if (ir == dst) continue;
                isymp[ir] = isymp[dst];
                isymp[dst] = cl;
                for (int& m : mo_irrep) {
                    if (m == ir + 1) m = dst + 1;
                    else if (m == dst + 1) m = ir + 1;
                }
                for (int a = 0; a < n; ++a) {
                    std::swap(vec[(size_t)ir * n + a], vec[(size_t)dst * n + a]);
                    std::swap(vs[(size_t)ir * n + a], vs[(size_t)dst * n + a]);
                }
            }
    }
    // ---- tmono.f:37-54: MO diagonal of h, then rotat2.f:177-203: degenerate levels
    std::vector<double> hc(nn);
    for (int i = 0; i < nvecf; ++i)
        for (int j = 1; j <= n; ++j) {
            double acc = 0.0;
            for (int k = 1; k <= n; ++k) acc = acc + c[(size_t)i * n + k - 1] * H(k, j);
            hc[(size_t)i * n + j - 1] = acc;
        }
    std::vector<double> eps(nvecf);
    for (int i = 0; i < nvecf; ++i) {      // the diagonal element of C^T h C, as tmono sums it
        double acc = 0.0;
        for (int k = 0; k < n; ++k) acc = acc + hc[(size_t)i * n + k] * c[(size_t)i * n + k];
        eps[i] = acc;
    }
    std::vector<int> lvl_first, lvl_size;
    for (int i = 0; i < nvecf;) {
        int d = 1;
        while (i + d < nvecf && std::fabs(eps[i] - eps[i + d]) < tdeg) ++d;
        lvl_first.push_back(i);
        lvl_size.push_back(d);
        i += d;
    }
    // ---- rotat2.f:210-323: irreps of every level, projection of degenerate levels
    std::vector<double> sc(nn);            // S c, one column per MO
    auto apply_s = [&]() {
        for (int i = 0; i < nvecf; ++i)
            for (int k = 1; k <= n; ++k) {
                double acc = 0.0;
                for (int l = 1; l <= n; ++l) acc = acc + c[(size_t)i * n + l - 1] * S(k, l);
                sc[(size_t)i * n + k - 1] = acc;
            }
    };
    apply_s();
    int err_tag = 0;
    std::vector<double> ov((size_t)nsym * n), proj;
    for (size_t lv = 0; lv < lvl_first.size() && !err_tag; ++lv) {
        const int i0 = lvl_first[lv], nd = lvl_size[lv];
        std::vector<int> hit;
        for (int ir = 0; ir < nsym; ++ir) {
            bool any = false;
            for (int j = 0; j < nd; ++j) {
                double acc = 0.0;
                for (int l = 0; l < n; ++l) acc = acc + vec[(size_t)ir * n + l] * sc[(size_t)(i0 + j) * n + l];
                ov[(size_t)ir * nd + j] = acc;
                any = any || std::fabs(acc) >= trec;
            }
            if (any) hit.push_back(ir);
        }
        const int nh = (int)hit.size();
        if (nh > nd || (nh < nd && nh != 1)) {
            err_tag = nh < nd ? 2 : 1;
        } else if (nh < nd) {                            // accidental degeneracy: no rotation
            for (int j = 0; j < nd; ++j) mo_irrep[i0 + j] = hit[0] + 1;
        } else {
            proj.assign((size_t)nd * n, 0.0);
            for (int t = 0; t < nd; ++t) {
                const int ir = hit[t];
                for (int l = 0; l < n; ++l) {
                    double acc = 0.0;
                    for (int j = 0; j < nd; ++j)
                        acc = acc + ov[(size_t)ir * nd + j] * c[(size_t)(i0 + j) * n + l];
                    proj[(size_t)t * n + l] = acc;
                }
                mo_irrep[i0 + t] = ir + 1;
            }
            std::memcpy(&c[(size_t)i0 * n], proj.data(), sizeof(double) * nd * n);
        }
    }
    // ---- rotat2.f:327-383: normalisation (tag -1) and orthogonality (tag 1)
    if (!err_tag) {
        apply_s();
        for (int i = 0; i < nvecf && !err_tag; ++i) {
            double* ci = &c[(size_t)i * n];
            double* si = &sc[(size_t)i * n];
            double nrm = 0.0;
            for (int k = 0; k < n; ++k) nrm = nrm + ci[k] * si[k];
            double r = std::sqrt(std::fabs(nrm));
            if (r < 1.e-6) { err_tag = -1; break; }
            r = 1.0 / r;
            for (int k = 0; k < n; ++k) {
                si[k] = si[k] * r;
                ci[k] = r * ci[k];
            }
        }
        for (int i = 1; i < nvecf && !err_tag; ++i)
            for (int j = 0; j < i; ++j) {
                double acc = 0.0;
                for (int k = 0; k < n; ++k) acc = acc + c[(size_t)j * n + k] * sc[(size_t)i * n + k];
                if (std::fabs(acc) > trec) { err_tag = 1; break; }
            }
    }
    if (err_tag) {
        if (tag) *tag = err_tag;
        return CIPSI_ERR_NOT_ORTHONORMAL;
    }
    // ---- rotat2.f:386-420: occupied MOs by irrep, then virtual MOs by irrep
    std::vector<double> out(c);
    {
        int n_occ = 0, n_virt = noc;
        for (int ir = 1; ir <= nsym; ++ir)
            for (int i = 0; i < nvecf; ++i) {
                if (mo_irrep[i] != ir) continue;
                const int k = i + 1 > noc ? n_virt++ : n_occ++;
                std::memcpy(&out[(size_t)k * n], &c[(size_t)i * n], sizeof(double) * n);
                itsyv[k] = ir;
            }
    }
    std::memcpy(c_out, out.data(), sizeof(double) * nn);
    // ---- vector.f:17-46
    int fill = 0;
    for (int ir = 0; ir < nsym; ++ir) {
        ind[ir] = fill + 1;
        int k = 0;
        for (int a = 0; a < n; ++a)
            if (vs[(size_t)ir * n + a] != 0) vs[(size_t)ir * n + a] = ++k;
        nbp[ir] = k;
        int m = 0;
        for (int i = 0; i < nvecf; ++i) {
            if (itsyv[i] != ir + 1) continue;
            ++m;
            for (int a = 0; a < n; ++a)
                if (vs[(size_t)ir * n + a] != 0) c_packed[fill++] = out[(size_t)i * n + a];
        }
        nbo[ir] = m;
    }
    ind[nsym] = fill + 1;
    std::memcpy(ivecsy, vs.data(), sizeof(int) * nsym * n);
    *nsymp = ncls;
    return CIPSI_OK;
}

}  // namespace

extern "C" int cipsi_ijkl_sym_prepare(int n_ao, int nsym, const int* its, const int* jsym, int nt,
                                      int nshell, const int* ktype, const int* kloc,
                                      const int* newsh, const int* iptr, const int* idtr,
                                      const int* iftr, const double* s_packed,
                                      const double* h_packed, const double* c_square, int nvecf,
                                      const int* ygel, int noc, double trec, double tdeg,
                                      double* c_out, int* itsyv, int* isymp, int* nsymp,
                                      int* ivecsy, int* nbp, int* nbo, int* ind, double* c_packed,
                                      int* tag) {
    (void)its;
    if (n_ao < 1 || nsym < 2 || nsym > 8 || nt < 1 || nshell < 1 || nvecf < 1 || nvecf > n_ao ||
        noc < 0 || !jsym || !ktype || !kloc || !newsh || !iptr || !idtr || !iftr || !h_packed ||
        !c_square || !c_out || !itsyv || !isymp || !nsymp || !ivecsy || !nbp || !nbo || !ind ||
        !c_packed)
        return CIPSI_ERR_ARG;
    const int n = n_ao;
    if (tag) *tag = 0;
    // ---- otodon.f:320-353: per irrep, AOs of its symmetry orbitals (signed SO number)
    std::vector<int> vs((size_t)nsym * n, 0);
    {
        int flat = 0;
        for (int ir = 0; ir < nsym; ++ir) {
            int so = 0;
            for (int sh = 0; sh < nshell; ++sh) {
                const int kt = ktype[sh];
                if (kt < 1 || kt > 4) return CIPSI_ERR_ARG;
                for (int comp = 0; comp < 2 * kt - 1; ++comp, ++flat) {
                    if (flat >= nsym * n) return CIPSI_ERR_ARG;
                    if (vs[flat] != 0) continue;
                    ++so;
                    for (int g = 0; g < nt; ++g) {
                        const int img_sh = newsh[sh + g * nshell] - 1;
                        if (img_sh < 0 || img_sh >= nshell) return CIPSI_ERR_ARG;
                        const int img = ir * n + kloc[img_sh] - 1 + comp;
                        if (img < ir * n || img >= (ir + 1) * n) return CIPSI_ERR_ARG;
                        int sgn = jsym[ir + g * nsym];
                        if (kt == 2) sgn *= iptr[comp + 3 * g];
                        else if (kt == 3) sgn *= idtr[comp + 6 * g];
                        else if (kt == 4) sgn *= iftr[comp + 10 * g];
                        const int v = so * sgn;
                        if (vs[img] != 0 && vs[img] != v) {       // the projection vanishes
                            for (int h = 0; h < nt; ++h)
                                vs[ir * n + kloc[newsh[sh + h * nshell] - 1] - 1 + comp] = 0;
                            --so;
                            break;
                        }
                        vs[img] = v;
                    }
                }
            }
        }
    }
    return rotat2_vector(n, nsym, vs, s_packed, h_packed, c_square, nvecf, ygel, noc, trec, tdeg,
                         c_out, itsyv, isymp, nsymp, ivecsy, nbp, nbo, ind, c_packed, tag);
}

// ------------------------------------------------------------------ linear groups
namespace {

// The classes ijkl gives the AOs of a linear molecule, by their angular character about
// the axis (otodon.f:37-39,62-144): sigma, pi_x, pi_y, delta_x, delta_y, phi_x, phi_y.
// kLinComp[class][ktype - 1]: the component of a shell of that type in the class (pure
// components, pshf's order: p = x, y, z; d = z^2, xz, yz, x^2 - y^2, xy), -1 if none.
const int kLinComp[7][4] = {{0, 2, 0, 0},   {-1, 0, 1, 1},   {-1, 1, 2, 2},  {-1, -1, 3, 3},
                            {-1, -1, 4, 4}, {-1, -1, -1, 5}, {-1, -1, -1, 6}};
// Parity of each class under x -> -x (bit 0) and y -> -y (bit 1): the irrep of the C2v
// subgroup the class reduces to.  The product table its is the subgroup's, with the
// labels the reference's tables itscv / itsdh give the subgroup irreps (BLOCKDATA1.f):
// C-inf-v by parity bits; D-inf-h by parity bits and g / u.
const int kLinBits[7] = {0, 1, 2, 0, 3, 1, 2};
const int kLinLabelCv[4] = {1, 2, 3, 5};
const int kLinLabelDh[4][2] = {{1, 2}, {3, 4}, {5, 6}, {9, 10}};
// The two symmetry transformations on the classes (isydec / isyded, ijkl.f:141-149):
// the rotation by 90 degrees about the axis (pi_x -> pi_y -> -pi_x, delta -> -delta,
// phi_x -> phi_y -> -phi_x) and the rotation by 45 degrees, kept on the sigma and delta
// classes only (delta_x -> delta_y -> -delta_x); 0: the class has no image.
const int kLinRot[2][7] = {{1, 3, -2, -4, -5, 7, -6}, {1, 0, 0, 5, -4, 0, 0}};

}  // namespace

extern "C" int cipsi_ijkl_sym_linear(const char* group, int n_ao, int nshell, const int* ktype,
                                     const int* kloc, const double* zsh, const int* newsh, int nt,
                                     int* nsym, int* its, int* ntrans, int* isydeg, int* so) {
    if (!group || n_ao < 1 || nshell < 1 || !ktype || !kloc || !zsh || !nsym || !its || !ntrans ||
        !isydeg || !so)
        return CIPSI_ERR_ARG;
    const bool dinfh = same_name(group, "DINFH");
    if (!dinfh && !same_name(group, "CINFV")) return CIPSI_ERR_ARG;
    if (dinfh && (!newsh || nt < 2)) return CIPSI_ERR_ARG;   // the image under sigma_h
    const int n = n_ao;
    bool has_p = false, has_d = false, has_f = false;
    for (int sh = 0; sh < nshell; ++sh) {
        const int kt = ktype[sh];
        if (kt < 1 || kt > 4 || kloc[sh] < 1 || kloc[sh] - 1 + 2 * kt - 1 > n) return CIPSI_ERR_ARG;
        if (dinfh && (newsh[sh + nshell] < 1 || newsh[sh + nshell] > nshell ||
                      ktype[newsh[sh + nshell] - 1] != kt))
            return CIPSI_ERR_ARG;
        has_p = has_p || kt == 2;
        has_d = has_d || kt == 3;
        has_f = has_f || kt == 4;
    }
    // otodon.f:47-60: one transformation for the p shells, one more for the d shells (the
    // presence of a higher shell implies the lower ones)
    if (has_f) has_d = true;
    if (has_d) has_p = true;
    const int ntr = (has_p ? 1 : 0) + (has_d ? 1 : 0);
    const int ns = dinfh ? (has_f ? 14 : 10) : (has_f ? 7 : 5);
    std::fill(so, so + (size_t)ns * n, 0);
    for (int ir = 0; ir < ns; ++ir) {
        const int cls = dinfh ? ir / 2 : ir;
        const bool gerade = dinfh && ir % 2 == 0;
        int* row = so + (size_t)ir * n;
        int count = 0;
        for (int sh = 0; sh < nshell; ++sh) {
            const int comp = kLinComp[cls][ktype[sh] - 1];
            if (comp < 0) continue;
            const int a = kloc[sh] - 1 + comp;
            if (!dinfh) {                                  // one AO per shell and class
                if (row[a] == 0) row[a] = ++count;
                continue;
            }
            // D-inf-h: the component plus or minus its image on the other atom; + when the
            // component's parity (s, d even; p, f odd) is the class's; an AO at the centre
            // enters only the classes of its own parity
            const bool odd = ktype[sh] % 2 == 0;
            const bool same = odd != gerade;
            const double z = std::fabs(zsh[sh]);
            if (!same && z < 1.e-8) continue;
            if (row[a] != 0) continue;
            row[a] = ++count;
            if (z > 1.e-8) {
                const int b = kloc[newsh[sh + nshell] - 1] - 1 + comp;
                if (row[b] == 0) row[b] = same ? count : -count;
            }
        }
    }
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j < ns; ++j) {
            int v;
            if (dinfh) {
                const int b = kLinBits[i / 2] ^ kLinBits[j / 2], par = (i ^ j) & 1;
                v = kLinLabelDh[b][par];
            } else {
                v = kLinLabelCv[kLinBits[i] ^ kLinBits[j]];
            }
            its[i + j * ns] = v;
        }
    for (int t = 0; t < ntr; ++t)
        for (int i = 0; i < ns; ++i) {
            int v;
            if (dinfh) {
                const int m = kLinRot[t][i / 2];
                v = m == 0 ? 0 : (m > 0 ? 1 : -1) * (2 * (std::abs(m) - 1) + i % 2 + 1);
            } else {
                v = kLinRot[t][i];
            }
            isydeg[i + t * ns] = v;
        }
    *nsym = ns;
    *ntrans = ntr;
    return CIPSI_OK;
}

extern "C" int cipsi_ijkl_sym_prepare_so(int n_ao, int nsym, const int* so, const double* s_packed,
                                         const double* h_packed, const double* c_square, int nvecf,
                                         const int* ygel, int noc, double trec, double tdeg,
                                         double* c_out, int* itsyv, int* isymp, int* nsymp,
                                         int* ivecsy, int* nbp, int* nbo, int* ind,
                                         double* c_packed, int* tag) {
    if (n_ao < 1 || nsym < 2 || nsym > 14 || nvecf < 1 || nvecf > n_ao || noc < 0 || !so ||
        !h_packed || !c_square || !c_out || !itsyv || !isymp || !nsymp || !ivecsy || !nbp ||
        !nbo || !ind || !c_packed)
        return CIPSI_ERR_ARG;
    if (tag) *tag = 0;
    std::vector<int> vs(so, so + (size_t)nsym * n_ao);
    return rotat2_vector(n_ao, nsym, vs, s_packed, h_packed, c_square, nvecf, ygel, noc, trec, tdeg,
                         c_out, itsyv, isymp, nsymp, ivecsy, nbp, nbo, ind, c_packed, tag);
}
