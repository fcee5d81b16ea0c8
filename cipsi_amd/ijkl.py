"""Python host mirror of the ijkl step, over the C ABI of libcipsi_ijkl.so.

Function names follow the reference routines they replace (CIPSI_sources/ijkl/):
  transd  -> unpack of the pshf label stream into the packed AO block (transd.f)
  tpqkld  -> first half-transform, packed AO block -> H[P_A, P_M]        (tpqkld.f)
  tijkld  -> second half-transform, H -> packed MO stream                (tijkld.f)
  itijkl  -> the whole transform from the label stream                   (itijkl.f)
  tmono   -> one-electron transform h_MO = C^T h C                       (tmono.f)
  transform -> the whole transform from a packed AO block

Arrays may be numpy (host entry points, synchronous) or torch CUDA tensors
(device entry points, asynchronous on torch's current stream).  Where the
reference stops with a message (ijkl.f:118-120, 304-316, rotat2.f:347-380) these
raise IjklError carrying the library status.  There is no CPU fallback: without the
built HIP library (or a gfx950 device) every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CIPSI_IJKL_LIBRARY selects another build of the library (same ABI)
LIB_PATH = os.environ.get("CIPSI_IJKL_LIBRARY") or os.path.join(_HERE, "libcipsi_ijkl.so")

PREC_FP64 = 0
PREC_SPLIT_FP16 = 1
_PREC = {"fp64": PREC_FP64, "split_fp16": PREC_SPLIT_FP16, PREC_FP64: PREC_FP64,
         PREC_SPLIT_FP16: PREC_SPLIT_FP16}

STATUS = {0: "CIPSI_OK", -1: "CIPSI_ERR_ARG", -2: "CIPSI_ERR_LABEL", -3: "CIPSI_ERR_NOMEM",
          -4: "CIPSI_ERR_HIP", -5: "CIPSI_ERR_NONFINITE", -6: "CIPSI_ERR_NOT_ORTHONORMAL",
          -7: "CIPSI_ERR_UNSUPPORTED", -8: "CIPSI_ERR_NODEVICE"}

# every symbol include/cipsi_ijkl.h declares
EXPORTS = [
    "cipsi_ijkl_create", "cipsi_ijkl_destroy", "cipsi_ijkl_strerror", "cipsi_ijkl_last_error",
    "cipsi_ijkl_set_stream", "cipsi_ijkl_reset_stream", "cipsi_ijkl_synchronize", "cipsi_ijkl_workspace_bytes",
    "cipsi_ijkl_release_workspace", "cipsi_ijkl_transform_packed", "cipsi_ijkl_transform_labels",
    "cipsi_ijkl_transform_packed_dev", "cipsi_ijkl_transform_labels_dev",
    "cipsi_ijkl_transform_packed_noalloc_dev", "cipsi_ijkl_unpack_labels_dev", "cipsi_ijkl_half1_dev", "cipsi_ijkl_half2_dev",
    "cipsi_ijkl_tmono", "cipsi_ijkl_enable_timing", "cipsi_ijkl_read_timing",
    "cipsi_ijkl_reset_timing", "cipsi_ijkl_autotune", "cipsi_ijkl_set_plan", "cipsi_ijkl_get_plan",
    "cipsi_ijkl_num_variants", "cipsi_ijkl_variant_info",
    "cipsi_ijkl_energy_grad", "cipsi_ijkl_energy_grad_dev",
    "cipsi_ijkl_batch_workspace_bytes", "cipsi_ijkl_batch_create", "cipsi_ijkl_batch_destroy",
    "cipsi_ijkl_batch_autotune", "cipsi_ijkl_batch_run", "cipsi_ijkl_batch_num_classes",
    "cipsi_ijkl_transform_batch_dev",
    "cipsi_ijkl_batch_class_info", "cipsi_ijkl_batch_set_class_plan",
    "cipsi_ijkl_gather_dev", "cipsi_ijkl_fock", "cipsi_ijkl_fock_dev",
    "cipsi_ijkl_hdig_dev", "cipsi_ijkl_hntd_dev", "cipsi_ijkl_sym_transform_labels",
    "cipsi_ijkl_sym_group", "cipsi_ijkl_sym_prepare", "cipsi_ijkl_sym_linear",
    "cipsi_ijkl_sym_prepare_so",
]


class SymDesc(ctypes.Structure):
    """cipsi_ijkl_sym: the COMMON state itijkl reads for a molecule with symmetry."""
    _fields_ = [("n_ao", ctypes.c_int), ("nsym", ctypes.c_int), ("nt", ctypes.c_int),
                ("nshell", ctypes.c_int)] + \
               [(f, ctypes.POINTER(ctypes.c_int)) for f in ("its", "isymp", "nbp", "nbo", "ind",
                                                            "ivecsy")] + \
               [("c", ctypes.POINTER(ctypes.c_double))] + \
               [(f, ctypes.POINTER(ctypes.c_int)) for f in ("ishell", "kloc", "ktyp", "numc",
                                                            "newsh", "iptr", "idtr", "iftr")] + \
               [("ntrsy", ctypes.c_int), ("isydeg", ctypes.POINTER(ctypes.c_int))]


class IjklError(RuntimeError):
    def __init__(self, status: int, what: str, detail: str = ""):
        self.status = status
        msg = f"{what}: {STATUS.get(status, status)}"
        if detail:
            msg += f" ({detail})"
        super().__init__(msg)


_lib = None
_c_int, _c_i64, _vp, _dp = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)


def load_library(path: str = LIB_PATH):
    """Load libcipsi_ijkl.so and declare its prototypes.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() "
                           "(make -C cipsi_amd/csrc)")
    # torch ships its own HIP runtime: when torch is installed it is loaded first, so the
    # library binds to that runtime instead of bringing up a second one (whose context
    # creation would then find no device)
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    L.cipsi_ijkl_create.argtypes = [ctypes.POINTER(_vp), _c_int]
    L.cipsi_ijkl_destroy.argtypes = [_vp]
    L.cipsi_ijkl_strerror.restype = ctypes.c_char_p
    L.cipsi_ijkl_strerror.argtypes = [_c_int]
    L.cipsi_ijkl_last_error.restype = ctypes.c_char_p
    L.cipsi_ijkl_last_error.argtypes = [_vp]
    L.cipsi_ijkl_set_stream.argtypes = [_vp, _vp]
    L.cipsi_ijkl_reset_stream.argtypes = [_vp]
    L.cipsi_ijkl_synchronize.argtypes = [_vp]
    L.cipsi_ijkl_workspace_bytes.restype = _c_i64
    L.cipsi_ijkl_workspace_bytes.argtypes = [_c_int, _c_int, _c_int, _c_int]
    L.cipsi_ijkl_release_workspace.argtypes = [_vp]
    L.cipsi_ijkl_transform_packed.argtypes = [_vp, _c_int, _c_int, _vp, _vp, _c_int, _vp, _c_int,
                                              _c_int, _vp]
    L.cipsi_ijkl_transform_labels.argtypes = [_vp, _c_int, _c_int, _vp, _vp, _c_i64, _vp, _c_int,
                                              _vp, _c_int, _c_int, _vp]
    L.cipsi_ijkl_transform_packed_dev.argtypes = [_vp, _c_int, _c_int, _vp, _vp, _c_int, _vp,
                                                  _c_int, _c_int]
    L.cipsi_ijkl_transform_labels_dev.argtypes = [_vp, _c_int, _c_int, _vp, _vp, _c_i64, _vp,
                                                  _c_int, _vp, _c_int, _c_int]
    L.cipsi_ijkl_transform_packed_noalloc_dev.argtypes = [_vp, _c_int, _c_int, _vp, _vp, _c_int,
                                                          _vp]
    L.cipsi_ijkl_unpack_labels_dev.argtypes = [_vp, _c_int, _vp, _vp, _c_i64, _vp]
    L.cipsi_ijkl_half1_dev.argtypes = [_vp, _c_int, _c_int, _vp, _vp, _c_int, _vp, _c_int, _c_int]
    L.cipsi_ijkl_half2_dev.argtypes = [_vp, _c_int, _c_int, _vp, _vp, _c_int, _vp, _c_int, _c_int]
    L.cipsi_ijkl_tmono.argtypes = [_vp, _c_int, _c_int, _vp, _vp, _c_int, _vp]
    L.cipsi_ijkl_enable_timing.argtypes = [_vp, _c_int]
    L.cipsi_ijkl_read_timing.argtypes = [_vp, ctypes.c_char_p, _dp, ctypes.POINTER(_c_i64)]
    L.cipsi_ijkl_reset_timing.argtypes = [_vp]
    _ip = ctypes.POINTER(_c_int)
    L.cipsi_ijkl_autotune.argtypes = [_vp, _c_int, _c_int, _c_int, _ip, _ip]
    L.cipsi_ijkl_set_plan.argtypes = [_vp, _c_int, _c_int, _c_int, _c_int, _c_int]
    L.cipsi_ijkl_get_plan.argtypes = [_vp, _c_int, _c_int, _c_int, _ip, _ip]
    L.cipsi_ijkl_num_variants.argtypes = []
    L.cipsi_ijkl_variant_info.argtypes = [_c_int, _ip, _ip, _ip]
    L.cipsi_ijkl_energy_grad.argtypes = [_vp, _c_int, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp,
                                         _vp, _vp]
    L.cipsi_ijkl_energy_grad_dev.argtypes = [_vp, _c_int, _c_int, _vp, _vp, _vp, _c_int, _vp,
                                             _vp, _vp, _vp]
    L.cipsi_ijkl_batch_workspace_bytes.restype = _c_i64
    L.cipsi_ijkl_batch_workspace_bytes.argtypes = [_c_int, _ip, _ip]
    L.cipsi_ijkl_batch_create.argtypes = [_vp, _c_int, _ip, _ip, ctypes.POINTER(_vp)]
    L.cipsi_ijkl_batch_destroy.argtypes = [_vp, _vp]
    L.cipsi_ijkl_batch_autotune.argtypes = [_vp, _vp, _vp, _vp, _ip, _vp]
    L.cipsi_ijkl_batch_run.argtypes = [_vp, _vp, _vp, _vp, _ip, _vp, _c_int]
    L.cipsi_ijkl_batch_num_classes.argtypes = [_vp]
    L.cipsi_ijkl_transform_batch_dev.argtypes = [_vp, _c_int, _ip, _ip, _vp, _vp, _ip, _vp,
                                                 _c_int, _c_int]
    L.cipsi_ijkl_batch_class_info.argtypes = [_vp, _c_int, _ip, _ip, _ip, _ip]
    L.cipsi_ijkl_batch_set_class_plan.argtypes = [_vp, _vp, _c_int, _c_int, _c_int]
    L.cipsi_ijkl_gather_dev.argtypes = [_vp, _c_int, _vp, _vp, _c_i64, _vp]
    L.cipsi_ijkl_fock.argtypes = [_vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]
    L.cipsi_ijkl_fock_dev.argtypes = [_vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]
    L.cipsi_ijkl_hdig_dev.argtypes = [_vp, _c_int, _c_int, _c_i64] + [_vp] * 10
    L.cipsi_ijkl_hntd_dev.argtypes = [_vp, _c_int, _c_int, _c_i64] + [_vp] * 7 + [_c_i64, _vp, _vp, _vp]
    _ipi = ctypes.POINTER(_c_int)
    L.cipsi_ijkl_sym_group.argtypes = [ctypes.c_char_p, _c_int, _c_int, _ipi, _ipi, _ipi]
    L.cipsi_ijkl_sym_prepare.argtypes = [_c_int, _c_int, _vp, _vp, _c_int, _c_int] + [_vp] * 6 + \
        [_vp, _vp, _vp, _c_int, _vp, _c_int, ctypes.c_double, ctypes.c_double] + [_vp] * 9 + [_vp]
    L.cipsi_ijkl_sym_linear.argtypes = [ctypes.c_char_p, _c_int, _c_int, _vp, _vp, _vp, _vp, _c_int,
                                        _ipi, _vp, _ipi, _vp, _vp]
    L.cipsi_ijkl_sym_prepare_so.argtypes = [_c_int, _c_int, _vp, _vp, _vp, _vp, _c_int, _vp, _c_int,
                                            ctypes.c_double, ctypes.c_double] + [_vp] * 9 + [_vp]
    _i16p = ctypes.POINTER(ctypes.c_int16)
    L.cipsi_ijkl_sym_transform_labels.argtypes = [_vp, ctypes.POINTER(SymDesc), _vp, _vp, _c_i64,
                                                  _vp, _c_i64, ctypes.POINTER(_c_i64), _vp, _c_int,
                                                  ctypes.POINTER(_c_int), _i16p, _i16p]
    _lib = L
    return L


def sym_group(group: str, naxis: int, nt: int):
    """(nsym, its, jsym) of a non-degenerate point group (cipsi_ijkl_sym_group,
    otodon.f:163-317); its is nsym x nsym, jsym nsym x nt (irrep x operation)."""
    L = load_library()
    ns = _c_int(0)
    its = np.zeros(64, dtype=np.int32)
    jsym = np.zeros(64, dtype=np.int32)
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(_c_int))
    rc = L.cipsi_ijkl_sym_group(group.encode().ljust(8), naxis, nt, ctypes.byref(ns), ip(its), ip(jsym))
    if rc != 0:
        raise IjklError(rc, "cipsi_ijkl_sym_group", L.cipsi_ijkl_strerror(rc).decode())
    k = ns.value
    return k, its[:k * k].reshape(k, k, order="F").copy(), jsym[:k * nt].reshape(k, nt, order="F").copy()


def sym_prepare(tables, s_packed, h_packed, c_square, noc, nvecf=None, ygel=None, trec=1e-5,
                tdeg=1e-6):
    """The MO preparation ijkl does before itijkl for a molecule with symmetry
    (cipsi_ijkl_sym_prepare: otodon + rotat2 + vector), on the host.  tables: norb, nsym,
    its, jsym (nsym x nt), nt, ktype, kloc, newsh (nshell x nt), iptr, idtr, iftr (1-based
    values).  s_packed None means S = 1.  Returns a dict with c (the ijom coefficients),
    itsyv, isymp, nsymp, ivecsy (nsym x n), nbp, nbo, ind, c_packed; raises IjklError
    (status CIPSI_ERR_NOT_ORTHONORMAL, .tag = rotat2's error tag) where the reference stops."""
    L = load_library()
    n, nsym, nt = int(tables["norb"]), int(tables["nsym"]), int(tables["nt"])
    nvecf = n if nvecf is None else nvecf
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    fort = lambda a: i32(np.asarray(a).T.ravel())
    its, jsym = fort(tables["its"]), fort(tables["jsym"])
    ktype, kloc = i32(tables["ktype"]), i32(tables["kloc"])
    newsh, iptr = fort(tables["newsh"]), fort(tables["iptr"])
    idtr, iftr = fort(tables["idtr"]), fort(tables["iftr"])
    sp = None if s_packed is None else np.ascontiguousarray(s_packed, dtype=np.float64)
    hp = np.ascontiguousarray(h_packed, dtype=np.float64)
    cs = np.asfortranarray(c_square, dtype=np.float64)
    yg = None if ygel is None else i32(ygel)
    c_out = np.zeros((n, n), order="F")
    itsyv = np.zeros(nvecf, dtype=np.int32)
    isymp = np.zeros(nsym, dtype=np.int32)
    nsymp = _c_int(0)
    ivecsy = np.zeros(nsym * n, dtype=np.int32)
    nbp = np.zeros(nsym, dtype=np.int32)
    nbo = np.zeros(nsym, dtype=np.int32)
    ind = np.zeros(nsym + 1, dtype=np.int32)
    cpk = np.zeros(n * n)
    tag = _c_int(0)
    pv = lambda a: None if a is None else _ptr(a)
    rc = L.cipsi_ijkl_sym_prepare(n, nsym, _ptr(its), _ptr(jsym), nt, ktype.size, _ptr(ktype),
                                  _ptr(kloc), _ptr(newsh), _ptr(iptr), _ptr(idtr), _ptr(iftr),
                                  pv(sp), _ptr(hp), _ptr(cs), nvecf, pv(yg), noc, trec, tdeg,
                                  _ptr(c_out), _ptr(itsyv), _ptr(isymp), ctypes.byref(nsymp),
                                  _ptr(ivecsy), _ptr(nbp), _ptr(nbo), _ptr(ind), _ptr(cpk),
                                  ctypes.byref(tag))
    if rc != 0:
        err = IjklError(rc, "cipsi_ijkl_sym_prepare", f"rotat2 error tag {tag.value}")
        err.tag = tag.value
        raise err
    return dict(c=c_out, itsyv=itsyv, isymp=isymp, nsymp=nsymp.value,
                ivecsy=ivecsy.reshape(nsym, n), nbp=nbp, nbo=nbo, ind=ind,
                c_packed=cpk[:int(ind[nsym]) - 1].copy())


def sym_linear(group: str, tables):
    """otodon's tables for the linear groups CINFV / DINFH (cipsi_ijkl_sym_linear,
    otodon.f:36-144).  tables: norb, ktype, kloc (per shell, 1-based), zsh (z of each
    shell's atom), newsh (nshell x nt, DINFH).  Returns (nsym, its (nsym x nsym), ntrans,
    isydeg (nsym x ntrans), so (nsym x norb))."""
    L = load_library()
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    n = int(tables["norb"])
    ktype, kloc = i32(tables["ktype"]), i32(tables["kloc"])
    zsh = np.ascontiguousarray(tables["zsh"], dtype=np.float64)
    newsh = i32(np.asarray(tables["newsh"]).T.ravel())
    nt = int(np.asarray(tables["newsh"]).shape[1])
    ns, ntr = _c_int(0), _c_int(0)
    its = np.zeros(14 * 14, dtype=np.int32)
    isydeg = np.zeros(14 * 2, dtype=np.int32)
    so = np.zeros(14 * n, dtype=np.int32)
    rc = L.cipsi_ijkl_sym_linear(group.encode().ljust(8), n, ktype.size, _ptr(ktype), _ptr(kloc),
                                 _ptr(zsh), _ptr(newsh), nt, ctypes.byref(ns), _ptr(its),
                                 ctypes.byref(ntr), _ptr(isydeg), _ptr(so))
    if rc != 0:
        raise IjklError(rc, "cipsi_ijkl_sym_linear", L.cipsi_ijkl_strerror(rc).decode())
    k, t = ns.value, ntr.value
    return (k, its[:k * k].reshape(k, k, order="F").copy(), t,
            isydeg[:k * t].reshape(k, t, order="F").copy(), so[:k * n].reshape(k, n).copy())


def sym_prepare_so(n_ao, nsym, so, s_packed, h_packed, c_square, noc, nvecf=None, ygel=None,
                   trec=1e-5, tdeg=1e-6):
    """rotat2 + vector for otodon's symmetry-orbital table so (nsym x n_ao), the linear
    groups' route (cipsi_ijkl_sym_prepare_so).  Returns what sym_prepare returns."""
    L = load_library()
    n = int(n_ao)
    nvecf = n if nvecf is None else nvecf
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    so = i32(np.asarray(so).ravel())
    sp = None if s_packed is None else np.ascontiguousarray(s_packed, dtype=np.float64)
    hp = np.ascontiguousarray(h_packed, dtype=np.float64)
    cs = np.asfortranarray(c_square, dtype=np.float64)
    yg = None if ygel is None else i32(ygel)
    c_out = np.zeros((n, n), order="F")
    itsyv = np.zeros(nvecf, dtype=np.int32)
    isymp = np.zeros(nsym, dtype=np.int32)
    nsymp = _c_int(0)
    ivecsy = np.zeros(nsym * n, dtype=np.int32)
    nbp = np.zeros(nsym, dtype=np.int32)
    nbo = np.zeros(nsym, dtype=np.int32)
    ind = np.zeros(nsym + 1, dtype=np.int32)
    cpk = np.zeros(n * n)
    tag = _c_int(0)
    pv = lambda a: None if a is None else _ptr(a)
    rc = L.cipsi_ijkl_sym_prepare_so(n, nsym, _ptr(so), pv(sp), _ptr(hp), _ptr(cs), nvecf, pv(yg),
                                     noc, trec, tdeg, _ptr(c_out), _ptr(itsyv), _ptr(isymp),
                                     ctypes.byref(nsymp), _ptr(ivecsy), _ptr(nbp), _ptr(nbo),
                                     _ptr(ind), _ptr(cpk), ctypes.byref(tag))
    if rc != 0:
        err = IjklError(rc, "cipsi_ijkl_sym_prepare_so", f"rotat2 error tag {tag.value}")
        err.tag = tag.value
        raise err
    return dict(c=c_out, itsyv=itsyv, isymp=isymp, nsymp=nsymp.value,
                ivecsy=ivecsy.reshape(nsym, n), nbp=nbp, nbo=nbo, ind=ind,
                c_packed=cpk[:int(ind[nsym]) - 1].copy())


def n_pairs(n: int) -> int:
    return n * (n + 1) // 2


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _ptr(x):
    if _is_torch(x):
        return ctypes.c_void_p(x.data_ptr())
    return ctypes.c_void_p(x.ctypes.data)


class Context:
    """Owns a library context (device workspace, stream, tile plans) on one GPU."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = _vp()
        rc = self.lib.cipsi_ijkl_create(ctypes.byref(h), device)
        if rc != 0:
            raise IjklError(rc, "cipsi_ijkl_create", self.lib.cipsi_ijkl_strerror(rc).decode())
        self.handle = h
        self.device = device

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cipsi_ijkl_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --------------------------------------------------------------- plumbing
    def _check(self, rc, what):
        if rc != 0:
            detail = self.lib.cipsi_ijkl_last_error(self.handle).decode(errors="replace")
            raise IjklError(rc, what, detail or self.lib.cipsi_ijkl_strerror(rc).decode())

    def _use_torch_stream(self):
        """Order library work with torch's current stream (handle 0 is the null stream)."""
        import torch
        self.lib.cipsi_ijkl_set_stream(self.handle, _vp(torch.cuda.current_stream().cuda_stream))

    def synchronize(self):
        self._check(self.lib.cipsi_ijkl_synchronize(self.handle), "cipsi_ijkl_synchronize")

    def release_workspace(self):
        self._check(self.lib.cipsi_ijkl_release_workspace(self.handle), "release_workspace")

    @staticmethod
    def _cmat(c):
        """(array, ldc): C as column-major N x M with leading dimension N."""
        if _is_torch(c):
            import torch
            if c.dtype != torch.float64 or not c.is_cuda:
                raise TypeError("C must be a float64 CUDA tensor")
            ct = c.t().contiguous()            # (M, N) row-major == (N, M) column-major
            return ct, c.shape[0]
        c = np.asfortranarray(c, dtype=np.float64)
        return c, c.shape[0]

    # --------------------------------------------------------------- operators
    def transform(self, ao_packed, c, precision="fp64", n_slices=0, return_err=False, out=None):
        """Packed AO block -> packed MO stream (the whole C1 itijkl path).  Host arrays
        may pass `out`, a float64 array of P_M(P_M+1)/2 elements reused across calls."""
        prec = _PREC[precision]
        cm, ldc = self._cmat(c)
        n, m = c.shape
        nout = n_pairs(n_pairs(m))
        if _is_torch(ao_packed):
            import torch
            if ao_packed.numel() != n_pairs(n_pairs(n)):
                raise ValueError("ao_packed has the wrong length")
            out = torch.empty(nout, dtype=torch.float64, device=ao_packed.device)
            self._use_torch_stream()
            self._check(self.lib.cipsi_ijkl_transform_packed_dev(
                self.handle, n, m, _ptr(ao_packed), _ptr(cm), ldc, _ptr(out), prec, n_slices),
                "transform_packed_dev")
            return out
        ao = np.ascontiguousarray(ao_packed, dtype=np.float64)
        if ao.size != n_pairs(n_pairs(n)):
            raise ValueError("ao_packed has the wrong length")
        if out is None:
            out = np.empty(nout)
        elif (not isinstance(out, np.ndarray) or out.dtype != np.float64 or out.size != nout
              or not out.flags.c_contiguous):
            raise ValueError("out must be a contiguous float64 array of P_M(P_M+1)/2 elements")
        err = ctypes.c_double(0.0)
        self.lib.cipsi_ijkl_reset_stream(self.handle)
        self._check(self.lib.cipsi_ijkl_transform_packed(
            self.handle, n, m, _ptr(ao), _ptr(cm), ldc, _ptr(out), prec, n_slices,
            ctypes.byref(err)), "transform_packed")
        return (out, err.value) if return_err else out

    def transform_noalloc(self, ao_packed, c, out=None):
        """The FP64 transform through cipsi_ijkl_transform_packed_noalloc_dev: float64 CUDA
        tensors; the library allocates nothing (its workspace is a static device array, so
        N_AO <= 48).  `out` (optional) receives the packed MO stream."""
        import torch
        cm, ldc = self._cmat(c)
        n, m = c.shape
        if ao_packed.numel() != n_pairs(n_pairs(n)):
            raise ValueError("ao_packed has the wrong length")
        nout = n_pairs(n_pairs(m))
        if out is None:
            out = torch.empty(nout, dtype=torch.float64, device=ao_packed.device)
        elif out.numel() != nout or out.dtype != torch.float64:
            raise ValueError("out has the wrong length or type")
        self._use_torch_stream()
        self._check(self.lib.cipsi_ijkl_transform_packed_noalloc_dev(
            self.handle, n, m, _ptr(ao_packed), _ptr(cm), ldc, _ptr(out)),
            "transform_packed_noalloc_dev")
        return out

    def sym_itijkl(self, tables, xx, ix):
        """itijkl for a molecule with point-group symmetry (cipsi_ijkl_sym_transform_labels).

        tables: dict of the COMMON state after vector.f (1-based, as the reference holds
        it): n_ao, nsym, nt, its (nsym x nsym), isymp, nbp, nbo, ind, ivecsy (nsym x n_ao),
        c (vector.f's packed coefficients), ishell, ktyp, numc (per AO), kloc (per shell),
        newsh (nshell x nt), iptr (3 x nt), idtr (6 x nt), iftr (10 x nt), and for the
        linear groups isydeg (nsym x ntrsy, the symmetry transformations).  xx, ix: the
        pshf skeleton label stream (numpy).  Returns (mo_stream, ipair (npaire x 2), nad,
        ktypsy) -- the f50 stream and the f25 directory of itijkl.f:208-213,291-296."""
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        fort = lambda a: i32(np.asarray(a).T.ravel())      # column-major
        keep = dict(its=fort(tables["its"]), isymp=i32(tables["isymp"]), nbp=i32(tables["nbp"]),
                    nbo=i32(tables["nbo"]), ind=i32(tables["ind"]),
                    ivecsy=i32(np.asarray(tables["ivecsy"]).ravel()),
                    ishell=i32(tables["ishell"]), kloc=i32(tables["kloc"]), ktyp=i32(tables["ktyp"]),
                    numc=i32(tables["numc"]), newsh=fort(tables["newsh"]), iptr=fort(tables["iptr"]),
                    idtr=fort(tables["idtr"]), iftr=fort(tables["iftr"]))
        isydeg = tables.get("isydeg")
        ntrsy = 0 if isydeg is None else int(np.asarray(isydeg).shape[1])
        keep["isydeg"] = fort(isydeg) if ntrsy else np.zeros(1, dtype=np.int32)
        cpk = np.ascontiguousarray(tables["c"], dtype=np.float64)
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        nshell = np.asarray(tables["newsh"]).shape[0]
        desc = SymDesc(int(tables["n_ao"]), int(tables["nsym"]), int(tables["nt"]), nshell,
                       *[ip(keep[k]) for k in ("its", "isymp", "nbp", "nbo", "ind", "ivecsy")],
                       cpk.ctypes.data_as(_dp),
                       *[ip(keep[k]) for k in ("ishell", "kloc", "ktyp", "numc", "newsh", "iptr",
                                               "idtr", "iftr")],
                       ntrsy, ip(keep["isydeg"]))
        xx = np.ascontiguousarray(xx, dtype=np.float64)
        ix = np.ascontiguousarray(ix, dtype=np.int32)
        nsym = int(tables["nsym"])
        nsyt = min(2000, (nsym * (nsym + 1) // 2) * (nsym * (nsym + 1) // 2 + 1) // 2)
        count = _c_i64(0)
        npaire = _c_int(0)
        max_pairs = 8192
        ipair = np.zeros(2 * max_pairs, dtype=np.int32)
        nad = np.zeros(nsyt, dtype=np.int16)
        ktypsy = np.zeros(nsyt, dtype=np.int16)
        i16 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int16))
        self.lib.cipsi_ijkl_reset_stream(self.handle)
        args = lambda out, cap: (self.handle, ctypes.byref(desc), _ptr(xx), _ptr(ix), xx.size, out,
                                 cap, ctypes.byref(count), _ptr(ipair), max_pairs,
                                 ctypes.byref(npaire), i16(nad), i16(ktypsy))
        self._check(self.lib.cipsi_ijkl_sym_transform_labels(*args(None, 0)), "sym_transform_labels")
        out = np.empty(max(count.value, 1))
        self._check(self.lib.cipsi_ijkl_sym_transform_labels(*args(_ptr(out), out.size)),
                    "sym_transform_labels")
        npr = npaire.value
        return out[:count.value], ipair[:2 * npr].reshape(npr, 2).copy(), nad, ktypsy

    def transform_batch(self, aos, cs, mos=None):
        """A whole scan in one call (cipsi_ijkl_transform_batch_dev): lists of float64 CUDA
        tensors, one packed AO block, C (N_AO x N_MO) and output per geometry.  The context
        keeps the plan of its last size list: a new list builds and autotunes one on this
        data, the same list again replays the captured graph on torch's current stream."""
        import torch
        ns = [int(c.shape[0]) for c in cs]
        ms = [int(c.shape[1]) for c in cs]
        if mos is None:
            mos = [torch.empty(n_pairs(n_pairs(m)), dtype=torch.float64, device=a.device)
                   for a, m in zip(aos, ms)]
        ao_p, c_p, ldc, mo_p, keep = _batch_arrays(ns, ms, aos, cs, mos)
        g = len(ns)
        self._use_torch_stream()
        self._check(self.lib.cipsi_ijkl_transform_batch_dev(
            self.handle, g, (ctypes.c_int * g)(*ns), (ctypes.c_int * g)(*ms), ao_p, c_p, ldc,
            mo_p, 0, 0), "transform_batch_dev")
        self._batch_keep = keep    # column-major copies of C stay alive until the next call
        return mos

    def itijkl(self, xx, ix, c, precision="fp64", n_slices=0, return_err=False, out=None):
        """pshf label stream (values, 32-bit labels) -> packed MO stream.  Host arrays may
        pass `out`, a float64 array of P_M(P_M+1)/2 elements reused across calls."""
        prec = _PREC[precision]
        cm, ldc = self._cmat(c)
        n, m = c.shape
        nout = n_pairs(n_pairs(m))
        if _is_torch(xx):
            import torch
            out = torch.empty(nout, dtype=torch.float64, device=xx.device)
            self._use_torch_stream()
            self._check(self.lib.cipsi_ijkl_transform_labels_dev(
                self.handle, n, m, _ptr(xx), _ptr(ix), xx.numel(), _ptr(cm), ldc, _ptr(out),
                prec, n_slices), "transform_labels_dev")
            return out
        xx = np.ascontiguousarray(xx, dtype=np.float64)
        ix = np.ascontiguousarray(ix, dtype=np.int32)
        if out is None:
            out = np.empty(nout)
        elif (not isinstance(out, np.ndarray) or out.dtype != np.float64 or out.size != nout
              or not out.flags.c_contiguous):
            raise ValueError("out must be a contiguous float64 array of P_M(P_M+1)/2 elements")
        err = ctypes.c_double(0.0)
        self.lib.cipsi_ijkl_reset_stream(self.handle)
        self._check(self.lib.cipsi_ijkl_transform_labels(
            self.handle, n, m, _ptr(xx), _ptr(ix), xx.size, _ptr(cm), ldc, _ptr(out), prec,
            n_slices, ctypes.byref(err)), "transform_labels")
        return (out, err.value) if return_err else out

    def transd(self, xx, ix, n_ao):
        """Label stream -> packed AO block (device tensors)."""
        import torch
        out = torch.empty(n_pairs(n_pairs(n_ao)), dtype=torch.float64, device=xx.device)
        self._use_torch_stream()
        self._check(self.lib.cipsi_ijkl_unpack_labels_dev(
            self.handle, n_ao, _ptr(xx), _ptr(ix), xx.numel(), _ptr(out)), "unpack_labels_dev")
        return out

    def tpqkld(self, ao_packed, c, precision="fp64", n_slices=0):
        """First half-transform on device tensors: H[P_A, P_M]."""
        import torch
        cm, ldc = self._cmat(c)
        n, m = c.shape
        h = torch.empty((n_pairs(n), n_pairs(m)), dtype=torch.float64, device=ao_packed.device)
        self._use_torch_stream()
        self._check(self.lib.cipsi_ijkl_half1_dev(
            self.handle, n, m, _ptr(ao_packed), _ptr(cm), ldc, _ptr(h), _PREC[precision],
            n_slices), "half1_dev")
        return h

    def tijkld(self, h, c, precision="fp64", n_slices=0):
        """Second half-transform on device tensors: H -> packed MO stream."""
        import torch
        cm, ldc = self._cmat(c)
        n, m = c.shape
        h = h.contiguous()
        out = torch.empty(n_pairs(n_pairs(m)), dtype=torch.float64, device=h.device)
        self._use_torch_stream()
        self._check(self.lib.cipsi_ijkl_half2_dev(
            self.handle, n, m, _ptr(h), _ptr(cm), ldc, _ptr(out), _PREC[precision], n_slices),
            "half2_dev")
        return out

    # --------------------------------------------------------------- tile plans
    def autotune(self, n_ao, n_mo, precision="fp64"):
        """Time every tile variant of both halves at this size; returns (half1, half2)."""
        v1, v2 = _c_int(-1), _c_int(-1)
        self._check(self.lib.cipsi_ijkl_autotune(self.handle, n_ao, n_mo, _PREC[precision],
                                                 ctypes.byref(v1), ctypes.byref(v2)), "autotune")
        return v1.value, v2.value

    def set_plan(self, n_ao, n_mo, half1=-1, half2=-1, precision="fp64"):
        self._check(self.lib.cipsi_ijkl_set_plan(self.handle, n_ao, n_mo, _PREC[precision],
                                                 half1, half2), "set_plan")

    def get_plan(self, n_ao, n_mo, precision="fp64"):
        v1, v2 = _c_int(-1), _c_int(-1)
        self._check(self.lib.cipsi_ijkl_get_plan(self.handle, n_ao, n_mo, _PREC[precision],
                                                 ctypes.byref(v1), ctypes.byref(v2)), "get_plan")
        return v1.value, v2.value

    # --------------------------------------------------------------- timing
    def enable_timing(self, on=True):
        self._check(self.lib.cipsi_ijkl_enable_timing(self.handle, int(on)), "enable_timing")

    def reset_timing(self):
        self._check(self.lib.cipsi_ijkl_reset_timing(self.handle), "reset_timing")

    def read_timing(self, stage):
        """(total milliseconds, launches) of one pipeline stage since the last reset."""
        ms = ctypes.c_double(0.0)
        cnt = _c_i64(0)
        self._check(self.lib.cipsi_ijkl_read_timing(self.handle, stage.encode(), ctypes.byref(ms),
                                                    ctypes.byref(cnt)), "read_timing")
        return ms.value, cnt.value

    def tmono(self, h_packed, c):
        cm, ldc = self._cmat(c)
        n, m = c.shape
        hp = np.ascontiguousarray(h_packed, dtype=np.float64)
        out = np.empty(n_pairs(m))
        self.lib.cipsi_ijkl_reset_stream(self.handle)
        self._check(self.lib.cipsi_ijkl_tmono(self.handle, n, m, _ptr(hp), _ptr(cm), ldc,
                                              _ptr(out)), "tmono")
        return out

    def energy_grad(self, ao_packed, h_packed, c, gamma, Gamma_packed, want_grad=True):
        """E(C) = sum h^MO gamma + 1/2 sum Gamma (pq|rs)_MO and dE/dC (N x M).

        gamma is M x M symmetric; Gamma_packed is the pair density packed like the MO
        stream.  Torch CUDA tensors run on torch's current stream and return tensors;
        numpy arrays use the synchronous host entry point.
        """
        n, m = c.shape
        if _is_torch(ao_packed):
            import torch
            dev = ao_packed.device
            cm, ldc = self._cmat(c)
            args = [x.contiguous() for x in (ao_packed, h_packed, gamma, Gamma_packed)]
            for x, size in zip(args, (n_pairs(n_pairs(n)), n_pairs(n), m * m,
                                      n_pairs(n_pairs(m)))):
                if x.dtype != torch.float64 or not x.is_cuda or x.numel() != size:
                    raise ValueError("energy_grad: inputs must be float64 CUDA tensors of the "
                                     "documented sizes")
            e = torch.empty(1, dtype=torch.float64, device=dev)
            g = torch.empty((m, n), dtype=torch.float64, device=dev) if want_grad else None
            self._use_torch_stream()
            self._check(self.lib.cipsi_ijkl_energy_grad_dev(
                self.handle, n, m, _ptr(args[0]), _ptr(args[1]), _ptr(cm), ldc, _ptr(args[2]),
                _ptr(args[3]), _ptr(e), _ptr(g) if want_grad else None), "energy_grad")
            return e[0], (g.t() if want_grad else None)
        cm, ldc = self._cmat(c)
        arrs = [np.ascontiguousarray(x, dtype=np.float64)
                for x in (ao_packed, h_packed, gamma, Gamma_packed)]
        for x, size in zip(arrs, (n_pairs(n_pairs(n)), n_pairs(n), m * m, n_pairs(n_pairs(m)))):
            if x.size != size:
                raise ValueError("energy_grad: input of the wrong size")
        e = ctypes.c_double()
        g = np.zeros((n, m), order="F") if want_grad else None
        self.lib.cipsi_ijkl_reset_stream(self.handle)
        self._check(self.lib.cipsi_ijkl_energy_grad(
            self.handle, n, m, _ptr(arrs[0]), _ptr(arrs[1]), _ptr(cm), ldc, _ptr(arrs[2]),
            _ptr(arrs[3]), ctypes.byref(e), _ptr(g) if want_grad else None), "energy_grad")
        return e.value, g

    def gather(self, mo_packed, n_mo, labels):
        """ai(i,j,k,l) for each label (1-based MO indices in the 8-bit fields of the label
        codec), gathered from the packed MO stream on the device (fock/ai.f:12-35).
        mo_packed: float64 CUDA tensor; labels: int32 CUDA tensor.  Returns float64."""
        import torch
        if not (_is_torch(mo_packed) and mo_packed.is_cuda and mo_packed.dtype == torch.float64
                and mo_packed.numel() == n_pairs(n_pairs(n_mo))):
            raise ValueError("gather: mo_packed must be a float64 CUDA tensor of P_M(P_M+1)/2")
        if not (_is_torch(labels) and labels.is_cuda and labels.dtype == torch.int32):
            raise ValueError("gather: labels must be an int32 CUDA tensor")
        mo_packed = mo_packed.contiguous()
        labels = labels.contiguous().view(-1)
        out = torch.empty(labels.numel(), dtype=torch.float64, device=labels.device)
        self._use_torch_stream()
        self._check(self.lib.cipsi_ijkl_gather_dev(self.handle, n_mo, _ptr(mo_packed),
                                                   _ptr(labels), labels.numel(), _ptr(out)),
                    "gather")
        return out

    def fock(self, mo_packed, h_mo_packed, n_mo, noc):
        """Closed-shell energy (without the nuclear term), Fock, J and K (packed) of
        fock.f:80-127 for noc doubly occupied MOs.  Torch CUDA tensors run on torch's
        current stream and return tensors (energy as a 0-d tensor); numpy arrays use the
        synchronous host entry point."""
        pm = n_pairs(n_mo)
        if _is_torch(mo_packed):
            import torch
            args = [x.contiguous() for x in (mo_packed, h_mo_packed)]
            for x, size in zip(args, (n_pairs(pm), pm)):
                if x.dtype != torch.float64 or not x.is_cuda or x.numel() != size:
                    raise ValueError("fock: inputs must be float64 CUDA tensors of the "
                                     "documented sizes")
            dev = mo_packed.device
            e = torch.empty(1, dtype=torch.float64, device=dev)
            f, j, k = (torch.empty(pm, dtype=torch.float64, device=dev) for _ in range(3))
            self._use_torch_stream()
            self._check(self.lib.cipsi_ijkl_fock_dev(self.handle, n_mo, noc, _ptr(args[0]),
                                                     _ptr(args[1]), _ptr(e), _ptr(f), _ptr(j),
                                                     _ptr(k)), "fock")
            return e[0], f, j, k
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (mo_packed, h_mo_packed)]
        if arrs[0].size != n_pairs(pm) or arrs[1].size != pm:
            raise ValueError("fock: input of the wrong size")
        e = ctypes.c_double()
        f, j, k = (np.zeros(pm) for _ in range(3))
        self.lib.cipsi_ijkl_reset_stream(self.handle)
        self._check(self.lib.cipsi_ijkl_fock(self.handle, n_mo, noc, _ptr(arrs[0]), _ptr(arrs[1]),
                                             ctypes.byref(e), _ptr(f), _ptr(j), _ptr(k)), "fock")
        return e.value, f, j, k

    def hdig(self, dets, n_mo, fock, jint, kint):
        """E(D) - E(reference) for every determinant (cip/hdig.f), float64 CUDA tensor."""
        import torch
        out = torch.empty(dets.ne.numel(), dtype=torch.float64, device=dets.ne.device)
        args = [x.contiguous() for x in (fock, jint, kint)]
        self._use_torch_stream()
        self._check(self.lib.cipsi_ijkl_hdig_dev(
            self.handle, n_mo, dets.ne.numel(), dets.trou.numel(), *dets._args(),
            *[_ptr(x) for x in args], _ptr(out)), "hdig")
        return out

    def hntd(self, dets, n_mo, pairs, mo_packed, fock):
        """<D_a|H|D_b> for each row (a, b) of pairs (int32 CUDA tensor, 0-based), as
        cip/hntd.f computes it; float64 CUDA tensor."""
        import torch
        pairs = pairs.contiguous().view(-1)
        if pairs.dtype != torch.int32 or pairs.numel() % 2:
            raise ValueError("hntd: pairs must be an int32 tensor of (a, b) rows")
        n_pairs = pairs.numel() // 2
        out = torch.empty(n_pairs, dtype=torch.float64, device=pairs.device)
        mo_packed, fock = mo_packed.contiguous(), fock.contiguous()
        self._use_torch_stream()
        self._check(self.lib.cipsi_ijkl_hntd_dev(
            self.handle, n_mo, dets.ne.numel(), dets.trou.numel(), *dets._args(), _ptr(pairs),
            n_pairs, _ptr(mo_packed), _ptr(fock), _ptr(out)), "hntd")
        return out


class Determinants:
    """cip's determinant arrays on the device (common /spo/, numbers 1-based): ne[d]
    excitations of the closed-shell reference, holes trou[nd[d]:nd[d]+ne[d]] and
    particles part[...] as spin-orbital numbers; spin-orbital s has orbital iorb[s-1] and
    spin ispin[s-1].  Accepts array-likes; keeps int32 CUDA tensors."""

    def __init__(self, ne, nd, trou, part, iorb, ispin, device="cuda"):
        import torch
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=device)
        self.ne, self.nd, self.trou, self.part = t(ne), t(nd), t(trou), t(part)
        self.iorb, self.ispin = t(iorb), t(ispin)

    def _args(self):
        return [_ptr(x) for x in (self.ne, self.nd, self.trou, self.part, self.iorb, self.ispin)]


class BatchPlan:
    """A scan of geometries transformed together (cipsi_ijkl_batch_*): one hipGraph
    launch of grouped kernels per run.  Inputs are lists of float64 CUDA
    tensors: packed AO blocks, C (N_AO x N_MO) and preallocated packed MO outputs."""

    def __init__(self, ctx: "Context", n_ao, n_mo):
        self.ctx = ctx
        self.lib = ctx.lib
        self.n = [int(x) for x in n_ao]
        self.m = [int(x) for x in n_mo]
        if len(self.n) != len(self.m):
            raise ValueError("n_ao and n_mo differ in length")
        na = (ctypes.c_int * len(self.n))(*self.n)
        nm = (ctypes.c_int * len(self.m))(*self.m)
        h = _vp()
        ctx._check(self.lib.cipsi_ijkl_batch_create(ctx.handle, len(self.n), na, nm,
                                                    ctypes.byref(h)), "batch_create")
        self.handle = h
        self._keep = None

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cipsi_ijkl_batch_destroy(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arrays(self, aos, cs, mos):
        return _batch_arrays(self.n, self.m, aos, cs, mos)

    def run(self, aos, cs, mos, use_graph=True):
        ao_p, c_p, ldc, mo_p, keep = self._arrays(aos, cs, mos)
        self.ctx._use_torch_stream()
        self.ctx._check(self.lib.cipsi_ijkl_batch_run(self.ctx.handle, self.handle, ao_p, c_p,
                                                      ldc, mo_p, int(bool(use_graph))),
                        "batch_run")
        self._keep = keep          # column-major copies of C stay alive until the next run
        return mos

    def autotune(self, aos, cs, mos):
        ao_p, c_p, ldc, mo_p, keep = self._arrays(aos, cs, mos)
        self.ctx._use_torch_stream()
        self.ctx._check(self.lib.cipsi_ijkl_batch_autotune(self.ctx.handle, self.handle, ao_p,
                                                           c_p, ldc, mo_p), "batch_autotune")
        self._keep = keep
        return self.classes()

    def classes(self):
        """[(row blocks, geometries, variant half 1, variant half 2)] per row class."""
        out = []
        for k in range(self.lib.cipsi_ijkl_batch_num_classes(self.handle)):
            v = [_c_int() for _ in range(4)]
            self.ctx._check(self.lib.cipsi_ijkl_batch_class_info(
                self.handle, k, *[ctypes.byref(x) for x in v]), "batch_class_info")
            out.append(tuple(x.value for x in v))
        return out

    def set_class_plan(self, k, var_half1, var_half2):
        self.ctx._check(self.lib.cipsi_ijkl_batch_set_class_plan(
            self.ctx.handle, self.handle, k, var_half1, var_half2), "batch_set_class_plan")


def _batch_arrays(ns, ms, aos, cs, mos):
    """Host arrays of device pointers for the batch entry points (and the column-major
    copies of C that must stay alive while the launch runs)."""
    g = len(ns)
    if not (len(aos) == len(cs) == len(mos) == g):
        raise ValueError("one AO block, C and output per geometry")
    ao_p = (_vp * g)()
    c_p = (_vp * g)()
    mo_p = (_vp * g)()
    ldc = (ctypes.c_int * g)()
    keep = []
    for i in range(g):
        n, m = ns[i], ms[i]
        if tuple(cs[i].shape) != (n, m):
            raise ValueError(f"geometry {i}: C must be {n} x {m}")
        if aos[i].numel() != n_pairs(n_pairs(n)) or mos[i].numel() != n_pairs(n_pairs(m)):
            raise ValueError(f"geometry {i}: AO block or output of the wrong length")
        for t in (aos[i], mos[i]):
            if not _is_torch(t) or not t.is_cuda or str(t.dtype) != "torch.float64" \
                    or not t.is_contiguous():
                raise TypeError("batch inputs must be contiguous float64 CUDA tensors")
        cm, ld = Context._cmat(cs[i])
        keep.append(cm)
        ao_p[i] = aos[i].data_ptr()
        c_p[i] = cm.data_ptr()
        mo_p[i] = mos[i].data_ptr()
        ldc[i] = ld
    return ao_p, c_p, ldc, mo_p, keep


def batch_workspace_bytes(n_ao, n_mo):
    g = len(n_ao)
    return load_library().cipsi_ijkl_batch_workspace_bytes(
        g, (ctypes.c_int * g)(*[int(x) for x in n_ao]), (ctypes.c_int * g)(*[int(x) for x in n_mo]))


def variants():
    """[(waves per work-group, panel depth, work-groups per CU hint)] of the built kernels."""
    L = load_library()
    out = []
    for v in range(L.cipsi_ijkl_num_variants()):
        w, k, b = _c_int(), _c_int(), _c_int()
        L.cipsi_ijkl_variant_info(v, ctypes.byref(w), ctypes.byref(k), ctypes.byref(b))
        out.append((w.value, k.value, b.value))
    return out


def workspace_bytes(n_ao, n_mo, precision="fp64", with_labels=False):
    return load_library().cipsi_ijkl_workspace_bytes(n_ao, n_mo, _PREC[precision], int(with_labels))
