"""Seeded synthetic inputs with point-group symmetry: a diatomic along z.

What pshf hands to ijkl for a molecule with symmetry is built here from a seed, so
that the reference ijkl and this implementation can be run on the same files:

* a basis of s, p and d shells (pure d, five components, as pshf writes them,
  CIPSI_sources/ijkl/rondof.f:51-76) on two atoms on the z axis;
* the group's operations in HONDO's order, as the info file carries them
  (rondof.f:30-38): the image shell of every shell (newsh) and the diagonal sign of
  every p and d component (ptr, dtr; rondof.f:100-118 keeps the diagonal only);
* AO integrals invariant under the group: (ij|kl) = sum_P B^P_ij B^P_kl where every
  factor B^P transforms as one irreducible representation, R_g B^P R_g^T = chi(g) B^P;
* the pshf "skeleton" label stream: one shell quartet per orbit of the group (the
  largest canonical image, pshf/jandk.f:152-189), its values multiplied by
  q4 = nt / |stabiliser| and divided by 2 for i==j, k==l, ij==kl (pshf/qout.f:71-77);
  transd regenerates the images (transd.f:125-244);
* symmetry-adapted MO coefficients: per irreducible representation, orthonormal
  combinations of its symmetry orbitals (AO combinations that otodon builds,
  otodon.f:320-353), columns interleaved across the representations as an SCF would
  order them by energy; a one-electron matrix h whose MO diagonal holds distinct
  orbital energies; S = 1.

Groups: C2v (heteronuclear: every operation keeps each atom) and D2h (homonuclear:
C2y, C2x, i and sigma_xy exchange the atoms).  The linear groups C-inf-v and D-inf-h, as
pshf writes them (pshf/ptgrp.f:416-428): the info file names the linear group but
carries the operations of the group pshf ran in, C1 for C-inf-v and Cs = (E, sigma_h)
for D-inf-h (sigma_h exchanges the atoms); ijkl then classifies the AOs itself by their
angular momentum about the axis (otodon.f:44-158).  Operations are in the order whose
characters are the rows of the group tables of CIPSI_sources/ijkl/BLOCKDATA1.f
(jsym4 / jsym8): C2v = (E, C2z, sigma_xz, sigma_yz), D2h = (E, C2z, C2y, C2x, i,
sigma_xy, sigma_xz, sigma_yz).

B is kept on a 2^-20 grid and the rank small, so every integral is an exact sum and the
data are bit-identical on every machine.
"""
from __future__ import annotations

import numpy as np

from . import synth

# axis signs (x, y, z) of each operation and whether it exchanges the two atoms
_OPS = {
    "C2v": [((1, 1, 1), False), ((-1, -1, 1), False), ((1, -1, 1), False), ((-1, 1, 1), False)],
    "D2h": [((1, 1, 1), False), ((-1, -1, 1), False), ((-1, 1, -1), True), ((1, -1, -1), True),
            ((-1, -1, -1), True), ((1, 1, -1), True), ((1, -1, 1), False), ((-1, 1, 1), False)],
}
# characters chi[irrep][op] (one-dimensional representations of these abelian groups)
_CHI = {
    "C2v": [[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, 1, -1], [1, -1, -1, 1]],
    "D2h": [[1, 1, 1, 1, 1, 1, 1, 1], [1, 1, -1, -1, 1, 1, -1, -1], [1, -1, 1, -1, 1, -1, 1, -1],
            [1, -1, -1, 1, 1, -1, -1, 1], [1, 1, 1, 1, -1, -1, -1, -1],
            [1, 1, -1, -1, -1, -1, 1, 1], [1, -1, 1, -1, -1, 1, -1, 1],
            [1, -1, -1, 1, -1, 1, 1, -1]],
}
# linear groups: the operations pshf runs in (C1; Cs with the plane normal to the axis)
_OPS["Cinfv"] = [((1, 1, 1), False)]
_OPS["Dinfh"] = [((1, 1, 1), False), ((1, 1, -1), True)]
LINEAR = ("Cinfv", "Dinfh")
GROUP_NAME = {"C2v": "CNV", "D2h": "DNH", "Cinfv": "CINFV", "Dinfh": "DINFH"}
# angular character of each component about the z axis: (m, 'c' cos / 's' sin / '0')
_MCHAR = {1: [(0, "0")], 2: [(1, "c"), (1, "s"), (0, "0")],
          3: [(0, "0"), (1, "c"), (1, "s"), (2, "c"), (2, "s")]}
# C-inf-v class of a component (otodon.f:118-132 with idinf2): sigma, pi_x, pi_y,
# delta_x, delta_y
_LCLASS = {(0, "0"): 0, (1, "c"): 1, (1, "s"): 2, (2, "c"): 3, (2, "s"): 4}
NFUNC = {1: 1, 2: 3, 3: 5}           # s, p, pure d (rondof.f:51-57 for pshf)
ATOM_Z = (1.25, -1.25)               # z coordinates of the two atoms (the info file's c)


def component_sign(ktype: int, comp: int, axes) -> int:
    """Sign of component `comp` of a shell under an operation with axis signs `axes`:
    p = (x, y, z); d = (z^2, xz, yz, x^2 - y^2, xy)."""
    sx, sy, sz = axes
    if ktype == 1:
        return 1
    if ktype == 2:
        return (sx, sy, sz)[comp]
    if ktype == 3:
        return (1, sx * sz, sy * sz, 1, sx * sy)[comp]
    raise ValueError("ktype")


class Molecule:
    """Shells on two atoms; shells of atom 1 first.  For D2h the two atoms carry the
    same shells (homonuclear)."""

    def __init__(self, group: str, shells_a, shells_b=None):
        if group not in _OPS:
            raise ValueError(group)
        self.group = group
        if shells_b is None or group in ("D2h", "Dinfh"):
            shells_b = list(shells_a)
        self.shells = [(0, k) for k in shells_a] + [(1, k) for k in shells_b]   # (atom, ktype)
        self.nshell = len(self.shells)
        self.kloc = []
        n = 0
        for _, kt in self.shells:
            self.kloc.append(n)
            n += NFUNC[kt]
        self.n_ao = n
        self.ops = _OPS[group]
        self.nt = len(self.ops)
        self.linear = group in LINEAR
        self.atom_z = ATOM_Z
        if self.linear:
            self.chi = None
            self.nsym = 5 if group == "Cinfv" else 10      # no f shells (otodon.f:66-67,113-114)
        else:
            self.chi = np.array(_CHI[group])
            self.nsym = self.chi.shape[0]
        na = len(shells_a)
        # image shell of every shell under every operation (0-based)
        self.newsh = np.zeros((self.nshell, self.nt), dtype=np.int64)
        for g, (_, swap) in enumerate(self.ops):
            for s, (atom, _) in enumerate(self.shells):
                if swap:
                    self.newsh[s, g] = s + na if atom == 0 else s - na
                else:
                    self.newsh[s, g] = s
        # AO tables: shell, component, type
        self.ishell = np.zeros(self.n_ao, dtype=np.int64)
        self.numc = np.zeros(self.n_ao, dtype=np.int64)
        self.ktyp = np.zeros(self.n_ao, dtype=np.int64)
        for s, (_, kt) in enumerate(self.shells):
            for c in range(NFUNC[kt]):
                a = self.kloc[s] + c
                self.ishell[a], self.numc[a], self.ktyp[a] = s, c, kt
        # signed permutation of the AOs under each operation: AO a -> sign * AO perm[a]
        self.perm = np.zeros((self.nt, self.n_ao), dtype=np.int64)
        self.sign = np.zeros((self.nt, self.n_ao))
        for g, (axes, _) in enumerate(self.ops):
            for a in range(self.n_ao):
                s = self.ishell[a]
                self.perm[g, a] = self.kloc[self.newsh[s, g]] + self.numc[a]
                self.sign[g, a] = component_sign(self.ktyp[a], self.numc[a], axes)

    def rep(self, g: int) -> np.ndarray:
        """R_g: R e_a = sign * e_perm[a]."""
        r = np.zeros((self.n_ao, self.n_ao))
        r[self.perm[g], np.arange(self.n_ao)] = self.sign[g]
        return r

    def ptr_diag(self):
        """(3, nt) and (6, nt) diagonal signs of the p and d transformation matrices."""
        p = np.array([[component_sign(2, c, axes) for (axes, _) in self.ops] for c in range(3)])
        d = np.array([[component_sign(3, c, axes) if c < 5 else 1 for (axes, _) in self.ops]
                      for c in range(6)])
        return p, d

    # ------------------------------------------------------------------ symmetry orbitals
    def symmetry_orbitals(self):
        """Per irreducible representation, the normalised projections sum_g chi(g) R_g e_a
        of the AOs, in AO order, skipping AOs already covered (otodon.f:320-353):
        a list over irreps of (n_so, N) arrays."""
        out = []
        for gam in range(self.nsym):
            used = np.zeros(self.n_ao, dtype=bool)
            rows = []
            for a in range(self.n_ao):
                if used[a]:
                    continue
                v = np.zeros(self.n_ao)
                for g in range(self.nt):
                    v[self.perm[g, a]] += self.chi[gam, g] * self.sign[g, a]
                imgs = np.unique(self.perm[:, a])
                used[imgs] = True
                if np.any(v != 0):
                    rows.append(v / np.sqrt(np.sum(v * v)))
            out.append(np.array(rows).reshape(-1, self.n_ao))
        return out


def eri_factors_sym(mol: Molecule, rank_per_irrep: int, decay: float, seed: int):
    """Factors B[P, pair0(i,j)]: rank_per_irrep of them per irreducible representation,
    each the projection (1/nt) sum_g chi(g) R_g B0 R_g^T of a seeded symmetric B0 on the
    2^-17 grid (so B is on the 2^-20 grid and every product sum is exact)."""
    rng = np.random.default_rng(seed)
    n = mol.n_ao
    a, b = np.tril_indices(n)
    env = np.exp(-np.abs(a - b) / float(decay))
    reps = [mol.rep(g) for g in range(mol.nt)]
    out = []
    for gam in range(mol.nsym):
        for _ in range(rank_per_irrep):
            z = np.clip(rng.standard_normal(a.size), -3.99, 3.99) * env
            z = np.rint(z * 2.0**17) * 2.0**-17
            b0 = np.zeros((n, n))
            b0[a, b] = z
            b0[b, a] = z
            acc = np.zeros((n, n))
            for g in range(mol.nt):
                acc += mol.chi[gam, g] * (reps[g] @ b0 @ reps[g].T)
            bm = acc / mol.nt
            out.append(bm[a, b])
    B = np.array(out)
    if B.shape[0] * 16.0 * 2.0**40 >= 2.0**53:
        raise ValueError("rank too large for exact integral sums")
    return B


def skeleton_labels(mol: Molecule, ao_packed: np.ndarray, cutoff: float = 0.0):
    """The pshf label stream of the symmetry-unique integrals (see the module notes).
    Returns (xx, ix) without the terminating zero label."""
    nt = mol.nt
    nsh = mol.nshell

    def canon(q):
        i, j, k, l = q
        if i < j:
            i, j = j, i
        if k < l:
            k, l = l, k
        if (i, j) < (k, l):
            i, j, k, l = k, l, i, j
        return (i, j, k, l)

    seen = set()
    xs, labs = [], []
    for i in range(nsh):
        for j in range(i + 1):
            for k in range(i + 1):
                for l in range(k + 1):
                    q = (i, j, k, l)
                    if canon(q) != q or q in seen:
                        continue
                    imgs = [canon(tuple(mol.newsh[s, g] for s in q)) for g in range(nt)]
                    rep = max(imgs)
                    n4 = sum(1 for x in imgs if x == rep)
                    for x in imgs:
                        seen.add(x)
                    q4 = nt / n4
                    si, sj, sk, sl = rep
                    emitted = set()
                    for ci in range(NFUNC[mol.shells[si][1]]):
                        for cj in range(NFUNC[mol.shells[sj][1]]):
                            for ck in range(NFUNC[mol.shells[sk][1]]):
                                for cl in range(NFUNC[mol.shells[sl][1]]):
                                    ao = canon((mol.kloc[si] + ci, mol.kloc[sj] + cj,
                                                mol.kloc[sk] + ck, mol.kloc[sl] + cl))
                                    if ao in emitted:
                                        continue
                                    emitted.add(ao)
                                    a1, a2, a3, a4 = ao
                                    v = ao_packed[synth.pair0(synth.pair0(a1, a2),
                                                              synth.pair0(a3, a4))]
                                    if v == 0.0 or abs(v) < cutoff:
                                        continue
                                    f = q4
                                    if a1 == a2:
                                        f *= 0.5
                                    if a3 == a4:
                                        f *= 0.5
                                    if a1 == a3 and a2 == a4:
                                        f *= 0.5
                                    xs.append(v * f)
                                    labs.append((a1 + 1, a2 + 1, a3 + 1, a4 + 1))
    from .formats import label_encode
    labs = np.array(labs, dtype=np.int64).reshape(-1, 4)
    return np.array(xs), label_encode(labs[:, 0], labs[:, 1], labs[:, 2], labs[:, 3])


def make_sym_case(group: str, shells_a, shells_b=None, rank_per_irrep: int = 3,
                  decay: float = 6.0, seed: int = 1, degenerate_pair: bool = False):
    """A whole symmetric input: molecule, AO integrals, skeleton stream, MOs, h.

    The MO columns of record 7 are the irreps' orbitals interleaved in a seeded random
    order; their orbital energies (the diagonal of C^T h C) are distinct and at least
    0.01 apart, except that with degenerate_pair two consecutive columns of different
    irreps share one energy (the case rotat2.f:172-323 rotates)."""
    mol = Molecule(group, shells_a, shells_b)
    n = mol.n_ao
    B = eri_factors_sym(mol, rank_per_irrep, decay, seed)
    ao = synth.eri_packed(B)
    xx, ix = skeleton_labels(mol, ao)
    rng = np.random.default_rng(seed + 1)
    sos = mol.symmetry_orbitals()
    cols, irr = [], []
    for gam, so in enumerate(sos):
        k = so.shape[0]
        if k == 0:
            continue
        q, r = np.linalg.qr(rng.standard_normal((k, k)))
        q = q * np.sign(np.diag(r))[None, :]
        for m in range(k):
            cols.append(q[:, m] @ so)         # exact mirror images on symmetry-related AOs
            irr.append(gam)
    order = rng.permutation(len(cols))
    c = np.asfortranarray(np.array([cols[o] for o in order]).T)
    irreps = np.array([irr[o] for o in order])
    e = np.sort(rng.uniform(-3.0, 3.0, size=n))
    e = e + 0.02 * np.arange(n)               # spacing >= 0.02
    if degenerate_pair:
        # two consecutive columns of different irreps with one energy
        for p in range(n - 1):
            if irreps[p] != irreps[p + 1]:
                e[p + 1] = e[p]
                break
    h = c @ np.diag(e) @ c.T
    h = 0.5 * (h + h.T)
    return dict(mol=mol, n_ao=n, B=B, ao_packed=ao, xx=xx, ix=ix, c_square=c,
                mo_irrep=irreps, h_packed=h[np.tril_indices(n)].copy(), energies=e,
                s_packed=np.eye(n)[np.tril_indices(n)].copy())


def check_invariance(mol: Molecule, ao_packed: np.ndarray) -> float:
    """max |(g i g j | g k g l) s - (ij|kl)| over all operations (0 for valid input)."""
    n = mol.n_ao
    full = synth.unpack_sym(ao_packed, n * (n + 1) // 2)
    worst = 0.0
    ia, ib = np.tril_indices(n)
    for g in range(mol.nt):
        p, s = mol.perm[g], mol.sign[g]
        pi = synth.pair0(np.maximum(p[ia], p[ib]), np.minimum(p[ia], p[ib]))
        sg = s[ia] * s[ib]
        img = full[np.ix_(pi, pi)] * np.outer(sg, sg)
        worst = max(worst, float(np.max(np.abs(img - full))))
    return worst


# ---------------------------------------------------------------------- linear groups
def linear_class_orbitals(mol: Molecule):
    """The symmetry orbitals of each of ijkl's classes for a linear molecule, in the
    order otodon numbers them (otodon.f:75-100, 118-132): for C-inf-v the AOs of one
    angular character (sigma, pi_x, pi_y, delta_x, delta_y), one per shell in shell
    order; for D-inf-h the classes (sigma, pi_x, pi_y, delta_x, delta_y) x (g, u), each
    orbital a shell's component on atom 1 plus or minus its image on atom 2, the sign
    + when the component's parity (s, d even; p odd) matches g.  A list over classes of
    (n_so, N) arrays, normalised."""
    n = mol.n_ao
    na = sum(1 for at, _ in mol.shells if at == 0)
    out = [[] for _ in range(mol.nsym)]
    for s, (atom, kt) in enumerate(mol.shells):
        for comp, mc in enumerate(_MCHAR[kt]):
            cls = _LCLASS[mc]
            a = mol.kloc[s] + comp
            if mol.group == "Cinfv":
                v = np.zeros(n)
                v[a] = 1.0
                out[cls].append(v)
                continue
            if atom != 0:
                continue
            b = mol.kloc[s + na] + comp
            odd = kt % 2 == 0                       # p shells
            for par in (0, 1):                      # 0 = g, 1 = u
                sign = 1.0 if (par == 0) != odd else -1.0
                v = np.zeros(n)
                v[a] = 1.0
                v[b] = sign
                out[2 * cls + par].append(v / np.sqrt(2.0))
    return [np.array(o).reshape(-1, n) for o in out]


def linear_ops(mol: Molecule):
    """Signed-permutation matrices the linear-group integrals are invariant under:
    the rotation by 90 degrees about z (x -> y, y -> -x), the reflection sigma_xz
    (y -> -y) and, for D-inf-h, the inversion (which exchanges the atoms)."""
    n = mol.n_ao
    na = sum(1 for at, _ in mol.shells if at == 0)
    rot = {1: [(0, 1)], 2: [(1, 1), (0, -1), (2, 1)], 3: [(0, 1), (2, 1), (1, -1), (3, -1), (4, -1)]}
    ref = {1: [1], 2: [1, -1, 1], 3: [1, 1, -1, 1, -1]}
    inv = {1: [1], 2: [-1, -1, -1], 3: [1, 1, 1, 1, 1]}
    mats = []
    r = np.zeros((n, n))
    f = np.zeros((n, n))
    for s, (atom, kt) in enumerate(mol.shells):
        for comp in range(NFUNC[kt]):
            a = mol.kloc[s] + comp
            tc, sg = rot[kt][comp]
            r[mol.kloc[s] + tc, a] = sg                   # R e_a = sg e_(image)
            f[a, a] = ref[kt][comp]
    mats += [r, f]
    if mol.group == "Dinfh":
        p = np.zeros((n, n))
        for s, (atom, kt) in enumerate(mol.shells):
            t = s + na if atom == 0 else s - na
            for comp in range(NFUNC[kt]):
                p[mol.kloc[t] + comp, mol.kloc[s] + comp] = inv[kt][comp]
        mats.append(p)
    return mats


def eri_factors_linear(mol: Molecule, rank: int, decay: float, seed: int):
    """Factors whose sum of squares sum_P B^P (x) B^P is invariant under every rotation
    about the axis and every reflection through it (and the inversion for D-inf-h),
    exactly: invariant m = 0 forms (sigma x sigma pairs, pi_x pi_x + pi_y pi_y,
    delta_x delta_x + delta_y delta_y) and doublets (B_c, B_s) that turn as
    (cos m phi, sin m phi): sigma x pi (m = 1), sigma x delta and pi x pi (m = 2).
    Coefficients are on the 2^-17 grid and each factor is a sum of symmetrised outer
    products of unit vectors, so B is on that grid and every integral sum is exact.
    For D-inf-h each factor is averaged with its inversion image (no division)."""
    rng = np.random.default_rng(seed)
    n = mol.n_ao
    comps = {}                      # (m, 'c'/'s'/'0') -> AO indices
    for s, (_, kt) in enumerate(mol.shells):
        for comp, mc in enumerate(_MCHAR[kt]):
            comps.setdefault(mc, []).append(mol.kloc[s] + comp)
    # partner of a cos component: the sin component of the same shell
    part = {}
    for s, (_, kt) in enumerate(mol.shells):
        for comp, (m, t) in enumerate(_MCHAR[kt]):
            if t == "c":
                part[mol.kloc[s] + comp] = mol.kloc[s] + _MCHAR[kt].index((m, "s"))

    def coef(a, b):
        z = np.clip(rng.standard_normal(), -3.99, 3.99) * np.exp(-abs(a - b) / float(decay))
        return np.rint(z * 2.0**17) * 2.0**-17

    def sym(u, v, w, mat):
        mat[u, v] += w
        mat[v, u] += w

    inv = linear_ops(mol)[2] if mol.group == "Dinfh" else None
    out = []

    def emit(mats):
        # D-inf-h: the gerade and the ungerade part of every factor (B + P B P^T,
        # B - P B P^T), so that blocks coupling g and u orbitals are not all zero
        for sgn in ((1.0, -1.0) if inv is not None else (0.0,)):
            for mm in mats:
                if inv is not None:
                    mm = mm + sgn * (inv @ mm @ inv.T)
                out.append(mm[np.tril_indices(n)])

    sig = comps.get((0, "0"), [])
    pic, dc = comps.get((1, "c"), []), comps.get((2, "c"), [])
    for _ in range(rank):
        # m = 0
        b0 = np.zeros((n, n))
        for i, a in enumerate(sig):
            for b in sig[: i + 1]:
                w = coef(a, b)
                if a == b:
                    b0[a, a] += w
                else:
                    sym(a, b, w, b0)
        for i, a in enumerate(pic + dc):
            for b in (pic + dc)[: i + 1]:
                if (a in pic) != (b in pic):
                    continue
                w = coef(a, b)
                if a == b:
                    b0[a, a] += w
                    b0[part[a], part[a]] += w
                else:
                    sym(a, b, w, b0)
                    sym(part[a], part[b], w, b0)
        emit([b0])
        # m = 1 doublets: sigma x pi
        if sig and pic:
            bc, bs = np.zeros((n, n)), np.zeros((n, n))
            for a in sig:
                for b in pic:
                    w = coef(a, b)
                    sym(a, b, w, bc)
                    sym(a, part[b], w, bs)
            emit([bc, bs])
        # m = 2 doublets: sigma x delta and pi x pi (cos cos - sin sin, cos sin + sin cos)
        if dc or len(pic) > 0:
            bc, bs = np.zeros((n, n)), np.zeros((n, n))
            for a in sig:
                for b in dc:
                    w = coef(a, b)
                    sym(a, b, w, bc)
                    sym(a, part[b], w, bs)
            for i, a in enumerate(pic):
                for b in pic[: i + 1]:
                    w = coef(a, b)
                    if a == b:
                        bc[a, a] += w
                        bc[part[a], part[a]] -= w
                        sym(a, part[a], w, bs)
                    else:
                        sym(a, b, w, bc)
                        sym(part[a], part[b], -w, bc)
                        sym(a, part[b], w, bs)
                        sym(part[a], b, w, bs)
            emit([bc, bs])
    B = np.array(out)
    if B.shape[0] * float(np.max(np.abs(B))) ** 2 * 2.0**34 >= 2.0**53:
        raise ValueError("rank too large for exact integral sums")
    return B


def make_linear_case(group: str, shells_a, shells_b=None, rank: int = 2, decay: float = 6.0,
                     seed: int = 1):
    """A whole linear-molecule input, as make_sym_case.  The MOs of each class are
    orthonormal combinations of its symmetry orbitals; the pi_x / pi_y (and delta_x /
    delta_y) partners of one orbital have the same coefficients on partner AOs (the
    rotation by 90 degrees maps one onto the other) and one energy, as an SCF gives
    them.  Columns are ordered by energy, x partner first."""
    mol = Molecule(group, shells_a, shells_b)
    n = mol.n_ao
    B = eri_factors_linear(mol, rank, decay, seed)
    ao = synth.eri_packed(B)
    xx, ix = skeleton_labels(mol, ao)
    rng = np.random.default_rng(seed + 1)
    sos = linear_class_orbitals(mol)
    if group == "Cinfv":
        fam = [(0, None), (1, 2), (3, 4)]
    else:
        fam = [(0, None), (1, None), (2, 4), (3, 5), (6, 8), (7, 9)]
    members = []                       # (energy, [(column, class), ...])
    for cx, cy in fam:
        k = sos[cx].shape[0]
        if k == 0:
            continue
        q, r = np.linalg.qr(rng.standard_normal((k, k)))
        q = q * np.sign(np.diag(r))[None, :]
        for m in range(k):
            cols = [(q[:, m] @ sos[cx], cx)]
            if cy is not None:
                cols.append((q[:, m] @ sos[cy], cy))
            members.append(cols)
    e = np.sort(rng.uniform(-3.0, 3.0, size=len(members))) + 0.02 * np.arange(len(members))
    order = rng.permutation(len(members))
    cols, irr, en = [], [], []
    for rank_e, mi in enumerate(order):
        for v, cls in members[mi]:
            cols.append(v)
            irr.append(cls)
            en.append(e[rank_e])
    c = np.asfortranarray(np.array(cols).T)
    en = np.array(en)
    h = c @ np.diag(en) @ c.T
    h = 0.5 * (h + h.T)
    return dict(mol=mol, n_ao=n, B=B, ao_packed=ao, xx=xx, ix=ix, c_square=c,
                mo_irrep=np.array(irr), h_packed=h[np.tril_indices(n)].copy(), energies=en,
                s_packed=np.eye(n)[np.tril_indices(n)].copy())


def check_invariance_linear(mol: Molecule, ao_packed: np.ndarray) -> float:
    """max |(Ri Rj|Rk Rl) - (ij|kl)| over the rotation by 90 degrees, sigma_xz, the
    inversion (D-inf-h) and, in floating point, the rotation by 45 degrees."""
    n = mol.n_ao
    full = synth.unpack_sym(ao_packed, n * (n + 1) // 2)
    mats = linear_ops(mol)
    c, s45 = np.cos(np.pi / 4), np.sin(np.pi / 4)
    r45 = np.zeros((n, n))
    for sh, (_, kt) in enumerate(mol.shells):
        k0 = mol.kloc[sh]
        if kt == 1:
            r45[k0, k0] = 1.0
        elif kt == 2:
            r45[k0:k0 + 2, k0:k0 + 2] = [[c, -s45], [s45, c]]
            r45[k0 + 2, k0 + 2] = 1.0
        else:
            r45[k0, k0] = 1.0
            r45[k0 + 1:k0 + 3, k0 + 1:k0 + 3] = [[c, -s45], [s45, c]]
            r45[k0 + 3:k0 + 5, k0 + 3:k0 + 5] = [[0.0, -1.0], [1.0, 0.0]]
    idx = np.arange(n)
    pq = synth.pair0(np.maximum.outer(idx, idx), np.minimum.outer(idx, idx))
    t4 = full[pq[:, :, None, None], pq[None, None, :, :]]          # (ij|kl), all index orders
    worst = 0.0
    for r in mats + [r45]:
        t4r = np.einsum("ai,bj,ck,dl,ijkl->abcd", r, r, r, r, t4, optimize=True)
        worst = max(worst, float(np.max(np.abs(t4r - t4))))
    return worst


def its_table(nsym: int) -> np.ndarray:
    """Product table of these abelian groups in this irrep order (BLOCKDATA1.f its2 /
    its4 / its8): the representations multiply as the XOR of their 0-based indices."""
    i = np.arange(nsym)
    return (np.bitwise_xor.outer(i, i) + 1).astype(np.int32)


def tables(mol: Molecule) -> dict:
    """The reference's view of the basis (1-based, rondof.f:77-118 and otodon's group
    tables): per shell ktype, kloc, newsh (nshell x nt); per AO ishell, kloc of its shell,
    ktyp, numc; diagonal signs iptr (3 x nt), idtr (6 x nt), iftr (10 x nt); characters
    jsym (nsym x nt) and the product table its (None for the linear groups, whose tables
    otodon builds from the basis); zsh, the z coordinate of each shell's atom."""
    pd, dd = mol.ptr_diag()
    return dict(
        norb=mol.n_ao, nsym=mol.nsym, nt=mol.nt,
        its=None if mol.linear else its_table(mol.nsym),
        jsym=None if mol.linear else mol.chi.astype(np.int32),
        zsh=np.array([ATOM_Z[atom] for atom, _ in mol.shells]),
        ktype=np.array([kt for _, kt in mol.shells], dtype=np.int32),
        kloc=np.array([mol.kloc[s] + 1 for s in range(mol.nshell)], dtype=np.int32),
        newsh=(mol.newsh + 1).astype(np.int32),
        iptr=pd.astype(np.int32), idtr=dd.astype(np.int32),
        iftr=np.ones((10, mol.nt), dtype=np.int32),
        ishell=(mol.ishell + 1).astype(np.int32),
        kloc_ao=np.array([mol.kloc[s] + 1 for s in mol.ishell], dtype=np.int32),
        ktyp=mol.ktyp.astype(np.int32), numc=mol.numc.astype(np.int32),
    )
