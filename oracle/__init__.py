"""ctypes wrapper of the CPU oracle (libijkl_oracle.so, built from ijkl_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or as the timed host-core baseline.
The product (cipsi_amd) never imports it.

Each function restates the reference routine named in ijkl_oracle.c
(CIPSI_sources/ijkl/transd.f, tpqkld.f, tijkld.f, tmono.f) for the C1 case, and in
ijkl_sym_oracle.c (otodon.f, rotat2.f, vector.f, itijkl.f, transd.f, tpqkld.f, tijkld.f)
for molecules with point-group symmetry.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libijkl_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.or_transd_c1.restype = ctypes.c_int64
        L.or_transd_c1.argtypes = [ctypes.c_int, _dp, _ip, ctypes.c_int64, _dp]
        for fn in (L.or_tpqkld_c1, L.or_tijkld_c1):
            fn.restype = None
            fn.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]
        for fn in (L.or_tpqkld_c1_range, L.or_tijkld_c1_range):
            fn.restype = None
            fn.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp,
                           ctypes.c_int64, ctypes.c_int64]
        for fn in (L.or_tpqkld_c1_omp_list, L.or_tijkld_c1_omp_list):
            fn.restype = None
            fn.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _lp,
                           ctypes.c_int64, ctypes.c_int]
        L.or_ijkl_c1.restype = ctypes.c_int
        L.or_ijkl_c1.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]
        L.or_tmono.restype = None
        L.or_tmono.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]
        L.or_energy.restype = ctypes.c_double
        L.or_energy.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.or_energy_grad.restype = None
        L.or_energy_grad.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int,
                                     _dp, _dp, _dp, _dp]
        L.or_fock.restype = ctypes.c_double
        L.or_fock.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        _ip32 = ctypes.POINTER(ctypes.c_int)
        L.or_hdig.restype = ctypes.c_double
        L.or_hdig.argtypes = [ctypes.c_int] + [_ip32] * 6 + [_dp, _dp, _dp]
        L.or_hntd.restype = ctypes.c_double
        L.or_hntd.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [_ip32] * 6 + [_dp, _dp]
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _cmat(c):
    c = np.asfortranarray(c, dtype=np.float64)
    return c, c.shape[0]


def n_pairs(n):
    return n * (n + 1) // 2


def transd(n_ao, xx, ix):
    """C1 unpack of a label stream into the packed AO block (transd.f:85-97,210-218,356-373)."""
    pa = n_pairs(n_ao)
    out = np.empty(n_pairs(pa))
    xx = _f64(xx)
    ix = np.ascontiguousarray(ix, dtype=np.int32)
    r = lib().or_transd_c1(n_ao, _d(xx), ix.ctypes.data_as(_ip), xx.size, _d(out))
    if r < 0:
        raise ValueError(f"label {-r - 1} maps outside the block")
    return out


def tpqkld(ao_packed, c):
    """First half-transform, pq-major H[npq, ilk] (tpqkld.f:25-104)."""
    c, ldc = _cmat(c)
    n, m = c.shape
    v = _f64(ao_packed)
    h = np.empty((n_pairs(n), n_pairs(m)))
    lib().or_tpqkld_c1(n, m, _d(v), _d(c), ldc, _d(h))
    return h


def tijkld(h, c):
    """Second half-transform to the packed MO stream (tijkld.f:28-88)."""
    c, ldc = _cmat(c)
    n, m = c.shape
    h = _f64(h)
    out = np.empty(n_pairs(n_pairs(m)))
    lib().or_tijkld_c1(n, m, _d(h), _d(c), ldc, _d(out))
    return out


def ijkl(ao_packed, c):
    """Whole C1 transform: tpqkld then tijkld."""
    c, ldc = _cmat(c)
    n, m = c.shape
    v = _f64(ao_packed)
    out = np.empty(n_pairs(n_pairs(m)))
    if lib().or_ijkl_c1(n, m, _d(v), _d(c), ldc, _d(out)) != 0:
        raise MemoryError("oracle allocation failed")
    return out


def tmono(h_packed, c):
    c, ldc = _cmat(c)
    n, m = c.shape
    hp = _f64(h_packed)
    out = np.empty(n_pairs(m))
    lib().or_tmono(n, m, _d(hp), _d(c), ldc, _d(out))
    return out


def energy(mo_packed, hmo_packed, gamma, Gamma_packed):
    m = gamma.shape[0]
    return lib().or_energy(0, m, _d(_f64(mo_packed)), _d(_f64(hmo_packed)), _d(_f64(gamma)),
                           _d(_f64(Gamma_packed)))


def fock(mo_packed, hmo_packed, n_mo, noc):
    """fock.f:80-127 on the packed MO stream: (energy without enuc, F, J, K), packed."""
    pm = n_pairs(n_mo)
    f, j, k = np.empty(pm), np.empty(pm), np.empty(pm)
    e = lib().or_fock(n_mo, noc, _d(_f64(mo_packed)), _d(_f64(hmo_packed)), _d(f), _d(j), _d(k))
    return e, f, j, k


class Determinants:
    """cip's determinant arrays (CIPSI_sources/cip, common /spo/), 1-based numbers:
    ne[d] excitations of the closed-shell reference, holes trou[nd[d]:nd[d]+ne[d]] and
    particles part[...] as spin-orbital numbers; spin-orbital s has orbital iorb[s-1]
    and spin ispin[s-1]."""

    def __init__(self, ne, nd, trou, part, iorb, ispin):
        as32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        self.ne, self.nd, self.trou, self.part = as32(ne), as32(nd), as32(trou), as32(part)
        self.iorb, self.ispin = as32(iorb), as32(ispin)

    def _args(self):
        ip = ctypes.POINTER(ctypes.c_int)
        return [a.ctypes.data_as(ip) for a in (self.ne, self.nd, self.trou, self.part,
                                                self.iorb, self.ispin)]


def hdig(dets, fock, jint, kint):
    """E(D) - E(reference) of every determinant (cip/hdig.f:40-80)."""
    fock, jint, kint = _f64(fock), _f64(jint), _f64(kint)
    a = dets._args()
    return np.array([lib().or_hdig(d, *a, _d(fock), _d(jint), _d(kint))
                     for d in range(dets.ne.size)])


def hntd(dets, pairs, norb, fock, mo_packed):
    """<D_a|H|D_b> for each pair (a, b) of 0-based determinant indices (cip/hntd.f:36-168)."""
    fock, mo = _f64(fock), _f64(mo_packed)
    a = dets._args()
    return np.array([lib().or_hntd(int(x), int(y), norb, *a, _d(fock), _d(mo))
                     for x, y in np.asarray(pairs).reshape(-1, 2)])


def energy_grad(ao_packed, h_packed, c, gamma, Gamma_packed):
    """E(C) and dE/dC (column-major N x M) in serial FP64."""
    c, ldc = _cmat(c)
    n, m = c.shape
    e = ctypes.c_double()
    g = np.zeros((n, m), order="F")
    lib().or_energy_grad(n, m, _d(_f64(ao_packed)), _d(_f64(h_packed)), _d(c), ldc,
                         _d(_f64(gamma)), _d(_f64(Gamma_packed)), ctypes.byref(e), _d(g))
    return e.value, g


def tpqkld_rows(ao_packed, c, npq_list, nthreads=1):
    """First half for the listed AO pairs only (rows of H; others left unset)."""
    c, ldc = _cmat(c)
    n, m = c.shape
    v = _f64(ao_packed)
    lst = np.ascontiguousarray(npq_list, dtype=np.int64)
    h = np.zeros((n_pairs(n), n_pairs(m)))
    lib().or_tpqkld_c1_omp_list(n, m, _d(v), _d(c), ldc, _d(h), lst.ctypes.data_as(_lp),
                                lst.size, nthreads)
    return h


def tijkld_rows(h, c, ilk_list, nthreads=1, out=None):
    """Second half for the listed MO pairs only."""
    c, ldc = _cmat(c)
    n, m = c.shape
    h = _f64(h)
    lst = np.ascontiguousarray(ilk_list, dtype=np.int64)
    if out is None:
        out = np.zeros(n_pairs(n_pairs(m)))
    lib().or_tijkld_c1_omp_list(n, m, _d(h), _d(c), ldc, _d(out), lst.ctypes.data_as(_lp),
                                lst.size, nthreads)
    return out


# ----------------------------------------------------------------------------- symmetry
class SymPrep:
    """State of the reference ijkl after rondof / otodon / rotat2 / vector for a molecule
    with point-group symmetry (ijkl_sym_oracle.c: or_sym_prepare)."""


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def sym_prepare(mol_tables, s_packed, h_packed, c_square, noc, nvecf=None, ygel=None,
                trec=1e-5, tdeg=1e-6):
    """otodon.f:320-353 + rotat2.f:50-423 + vector.f:17-46.  mol_tables: dict with norb,
    nsym, its (nsym x nsym), jsym (nsym x nt), nt, ktype, kloc (1-based), newsh (nshell x
    nt, 1-based), iptr (3 x nt), idtr (6 x nt), iftr (10 x nt).  Returns a SymPrep, or
    raises ValueError(tag) where rotat2 stops with an error tag."""
    L = lib()
    ip = ctypes.POINTER(ctypes.c_int)
    f = L.or_sym_prepare
    f.restype = ctypes.c_int
    t = mol_tables
    n, nsym, nt = t["norb"], t["nsym"], t["nt"]
    nshell = len(t["ktype"])
    nvecf = n if nvecf is None else nvecf
    yg = np.zeros(n, dtype=np.uint8) if ygel is None else np.ascontiguousarray(ygel, np.uint8)
    its = _i32(np.asarray(t["its"]).T.ravel())          # column-major its(i,j)
    jsym = _i32(np.asarray(t["jsym"]).T.ravel())
    newsh = _i32(np.asarray(t["newsh"]).T.ravel())
    iptr = _i32(np.asarray(t["iptr"]).T.ravel())
    idtr = _i32(np.asarray(t["idtr"]).T.ravel())
    iftr = _i32(np.asarray(t["iftr"]).T.ravel())
    ktype, kloc = _i32(t["ktype"]), _i32(t["kloc"])
    c_in = np.asfortranarray(c_square, dtype=np.float64)
    c_out = np.zeros((n, n), order="F")
    itsyv = np.zeros(n, dtype=np.int32)
    isymp = np.zeros(nsym, dtype=np.int32)
    nsymp = ctypes.c_int()
    ivecsy = np.zeros(nsym * n, dtype=np.int32)
    nbp = np.zeros(nsym, dtype=np.int32)
    nbo = np.zeros(nsym, dtype=np.int32)
    ind = np.zeros(nsym + 1, dtype=np.int32)
    cpk = np.zeros(n * n)
    p = lambda a: a.ctypes.data_as(ip)
    rc = f(n, nsym, p(its), p(jsym), nt, nshell, p(ktype), p(kloc), p(newsh), p(iptr), p(idtr),
           p(iftr), _d(_f64(s_packed)), _d(_f64(h_packed)), _d(c_in), nvecf,
           yg.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), noc, ctypes.c_double(trec),
           ctypes.c_double(tdeg), _d(c_out), p(itsyv), p(isymp), ctypes.byref(nsymp), p(ivecsy),
           p(nbp), p(nbo), p(ind), _d(cpk))
    if rc != 0:
        raise ValueError(rc)
    r = SymPrep()
    r.norb, r.nsym, r.nt, r.nvec = n, nsym, nt, nvecf
    r.its = np.asarray(t["its"], dtype=np.int32)
    r.c = c_out
    r.itsyv = itsyv[:nvecf].copy()
    r.isymp = isymp
    r.nsymp = nsymp.value
    r.ivecsy = ivecsy.reshape(nsym, n)
    r.nbp, r.nbo, r.ind = nbp, nbo, ind
    r.c_packed = cpk[:int(ind[nsym]) - 1].copy()
    return r


def sym_linear(group, mol_tables):
    """otodon.f:36-144 for CINFV / DINFH.  mol_tables: dict with norb, ktype, kloc (per
    shell, 1-based), zsh (z of each shell's atom) and, for DINFH, newsh (nshell x nt).
    Returns (nsym, its (nsym x nsym), ntrans, isydeg (nsym x ntrans), so (nsym x N))."""
    L = lib()
    ip = ctypes.POINTER(ctypes.c_int)
    f = L.or_sym_linear
    f.restype = ctypes.c_int
    t = mol_tables
    n = t["norb"]
    nshell = len(t["ktype"])
    dinfh = group.upper().strip() == "DINFH"
    if not dinfh and group.upper().strip() != "CINFV":
        raise ValueError(group)
    newsh2 = _i32(np.asarray(t["newsh"])[:, 1] if dinfh else np.zeros(nshell))
    its = np.zeros(14 * 14, dtype=np.int32)
    isydeg = np.zeros(14 * 2, dtype=np.int32)
    so = np.zeros(14 * n, dtype=np.int32)
    nsym, ntrans = ctypes.c_int(), ctypes.c_int()
    p = lambda a: a.ctypes.data_as(ip)
    f(int(dinfh), n, nshell, p(_i32(t["ktype"])), p(_i32(t["kloc"])), _d(_f64(t["zsh"])), p(newsh2),
      ctypes.byref(nsym), p(its), ctypes.byref(ntrans), p(isydeg), p(so))
    ns, nt = nsym.value, ntrans.value
    return (ns, its[:ns * ns].reshape(ns, ns).T.copy(), nt,
            isydeg[:ns * nt].reshape(nt, ns).T.copy(), so[:ns * n].reshape(ns, n).copy())


def sym_prepare_so(norb, nsym, its, so, s_packed, h_packed, c_square, noc, nvecf=None,
                   ygel=None, trec=1e-5, tdeg=1e-6):
    """rotat2.f:50-423 + vector.f:17-46 for the symmetry-orbital table so (nsym x N) that
    sym_linear returns.  Returns a SymPrep as sym_prepare."""
    L = lib()
    ip = ctypes.POINTER(ctypes.c_int)
    f = L.or_sym_prepare_so
    f.restype = ctypes.c_int
    n = norb
    nvecf = n if nvecf is None else nvecf
    yg = np.zeros(n, dtype=np.uint8) if ygel is None else np.ascontiguousarray(ygel, np.uint8)
    c_in = np.asfortranarray(c_square, dtype=np.float64)
    c_out = np.zeros((n, n), order="F")
    itsyv = np.zeros(n, dtype=np.int32)
    isymp = np.zeros(nsym, dtype=np.int32)
    nsymp = ctypes.c_int()
    ivecsy = np.zeros(nsym * n, dtype=np.int32)
    nbp = np.zeros(nsym, dtype=np.int32)
    nbo = np.zeros(nsym, dtype=np.int32)
    ind = np.zeros(nsym + 1, dtype=np.int32)
    cpk = np.zeros(n * n)
    p = lambda a: a.ctypes.data_as(ip)
    rc = f(n, nsym, p(_i32(np.asarray(so).ravel())), _d(_f64(s_packed)), _d(_f64(h_packed)),
           _d(c_in), nvecf, yg.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), noc,
           ctypes.c_double(trec), ctypes.c_double(tdeg), _d(c_out), p(itsyv), p(isymp),
           ctypes.byref(nsymp), p(ivecsy), p(nbp), p(nbo), p(ind), _d(cpk))
    if rc != 0:
        raise ValueError(rc)
    r = SymPrep()
    r.norb, r.nsym, r.nvec = n, nsym, nvecf
    r.its = np.asarray(its, dtype=np.int32)
    r.c = c_out
    r.itsyv = itsyv[:nvecf].copy()
    r.isymp = isymp
    r.nsymp = nsymp.value
    r.ivecsy = ivecsy.reshape(nsym, n)
    r.nbp, r.nbo, r.ind = nbp, nbo, ind
    r.c_packed = cpk[:int(ind[nsym]) - 1].copy()
    return r


def sym_itijkl(prep, ao_tables, xx, ix, out_cap=None, isydeg=None):
    """itijkl.f:42-298 with transd / tpqkld / tijkld in their symmetric forms.  ao_tables:
    dict with ishell, kloc, ktyp, numc (per AO, 1-based shell and kloc), newsh (nshell x nt,
    1-based), iptr, idtr, iftr.  isydeg: the symmetry transformations (nsym x ntrsy) of
    the linear groups, or None.  Returns (mo_stream, ipair (npaire x 2), nad, ktypsy)."""
    L = lib()
    ip = ctypes.POINTER(ctypes.c_int)
    f = L.or_sym_itijkl
    f.restype = ctypes.c_int
    t = ao_tables
    n, nsym = prep.norb, prep.nsym
    nshell, nt = np.asarray(t["newsh"]).shape
    if isydeg is None:
        isydeg = np.zeros((nsym, 0), dtype=np.int32)
    ntrsy = int(np.asarray(isydeg).shape[1])
    isy = _i32(np.asarray(isydeg).T.ravel()) if ntrsy else np.zeros(1, dtype=np.int32)
    its = _i32(prep.its.T.ravel())
    newsh = _i32(np.asarray(t["newsh"]).T.ravel())
    iptr = _i32(np.asarray(t["iptr"]).T.ravel())
    idtr = _i32(np.asarray(t["idtr"]).T.ravel())
    iftr = _i32(np.asarray(t["iftr"]).T.ravel())
    ishell, kloc, ktyp, numc = (_i32(t[k]) for k in ("ishell", "kloc", "ktyp", "numc"))
    ivecsy = _i32(prep.ivecsy.ravel())
    if out_cap is None:
        pm = n * (n + 1) // 2
        out_cap = pm * (pm + 1) // 2
    out = np.zeros(out_cap)
    nout = ctypes.c_int64()
    pair_cap = 4096
    ipair = np.zeros(2 * pair_cap, dtype=np.int32)
    npaire = ctypes.c_int()
    nsyt = min(2000, (nsym * (nsym + 1) // 2) * (nsym * (nsym + 1) // 2 + 1) // 2)
    nad = np.zeros(nsyt, dtype=np.int32)
    ktypsy = np.zeros(nsyt, dtype=np.int32)
    xx = _f64(xx)
    ix = np.ascontiguousarray(ix, dtype=np.int32)
    p = lambda a: a.ctypes.data_as(ip)
    rc = f(n, nsym, p(its), p(_i32(prep.isymp)), p(_i32(prep.nbp)), p(_i32(prep.nbo)),
           p(_i32(prep.ind)), p(ivecsy), _d(_f64(prep.c_packed)), nt, nshell, p(ishell), p(kloc),
           p(ktyp), p(numc), p(newsh), p(iptr), p(idtr), p(iftr), ntrsy, p(isy), _d(xx),
           ix.ctypes.data_as(_ip), xx.size, _d(out), out.size, ctypes.byref(nout), p(ipair),
           pair_cap, ctypes.byref(npaire), p(nad), p(ktypsy))
    if rc != 0:
        raise RuntimeError("or_sym_itijkl failed")
    npr = npaire.value
    return out[:nout.value].copy(), ipair[:2 * npr].reshape(npr, 2).copy(), nad, ktypsy
