"""File formats of the ijkl step (Fortran unformatted sequential files, C1 case).

Readers and writers for the files ijkl reads and writes, so that synthetic
inputs can be handed to either implementation and their outputs compared:

  info (unit 2)  written by pshf  CIPSI_sources/pshf/daclos.f:41-59, read by
                 ijkl.f:100-101, rondof.f:87-107, rotat2.f:33-38, tmono.f:19-22
  pqrs (unit 8)  written by pshf  pshf/qout.f:71-88, read by itijkl.f:57-68
  f50  (unit 50) written by tijkld.f:68-88 and tmono.f:56
  f25  (unit 25) written by ijkl.f:249-253,319 and itijkl.f:294-296
  ijom (unit 10) written by ijkl.f:223-240

Fortran sequential unformatted records are framed by a 4-byte little-endian
byte count before and after the payload.
"""
from __future__ import annotations

import struct

import numpy as np

LSIZE = 8192           # record length of the pqrs / f50 / scratch buffers (pshf.prm:5)


# --------------------------------------------------------------------------- records
def write_record(f, payload: bytes) -> None:
    n = len(payload)
    if n >= 2**31:
        raise ValueError("record too long for a 4-byte marker")
    f.write(struct.pack("<i", n))
    f.write(payload)
    f.write(struct.pack("<i", n))


def read_records(path) -> list[bytes]:
    recs = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        if n < 0:
            raise ValueError("sub-records are not supported")
        payload = data[pos + 4:pos + 4 + n]
        (n2,) = struct.unpack_from("<i", data, pos + 4 + n)
        if n2 != n:
            raise ValueError(f"corrupt record at byte {pos}")
        recs.append(payload)
        pos += 8 + n
    return recs


# --------------------------------------------------------------------------- labels
def label_encode(i, j, k, l):
    """pshf/iword_func.f:2-14: i<<24 | j<<16 | k<<8 | l (1-based AO indices)."""
    i = np.asarray(i, dtype=np.int64)
    v = (i << 24) | (np.asarray(j, dtype=np.int64) << 16) | (np.asarray(k, dtype=np.int64) << 8) \
        | np.asarray(l, dtype=np.int64)
    return v.astype(np.uint32).view(np.int32)


def label_decode(labels):
    """CIPSI_sources/ijkl/iword_func.f:2-52: the four 8-bit fields."""
    u = np.asarray(labels).astype(np.int32).view(np.uint32).astype(np.int64)
    return (u >> 24) & 255, (u >> 16) & 255, (u >> 8) & 255, u & 255


def labels_from_packed(ao_packed: np.ndarray, n_ao: int, cutoff: float = 0.0):
    """The pshf label stream for a packed AO block: canonical labels i>=j, k>=l,
    (ij)>=(kl), values divided by 2 for i==j, k==l and ij==kl (pshf/qout.f:52-77),
    entries with |value| < cutoff omitted (qout.f:47).  Returns (xx, ix) without
    the terminating zero label."""
    if n_ao > 255:
        raise ValueError("8-bit label fields cap N_AO at 255")
    pa = n_ao * (n_ao + 1) // 2
    ia, ib = np.tril_indices(n_ao)               # pair index -> (i, j), 0-based
    ab, cd = np.tril_indices(pa)                 # packed index -> (ij, kl)
    i, j = ia[ab], ib[ab]
    k, l = ia[cd], ib[cd]
    val = ao_packed.astype(np.float64).copy()
    keep = np.abs(val) >= cutoff if cutoff > 0 else np.ones(val.size, dtype=bool)
    scale = np.ones(val.size)
    scale[i == j] *= 0.5
    scale[k == l] *= 0.5
    scale[ab == cd] *= 0.5
    xx = (val * scale)[keep]
    ix = label_encode(i[keep] + 1, j[keep] + 1, k[keep] + 1, l[keep] + 1)
    return xx, ix


def write_pqrs(path, xx: np.ndarray, ix: np.ndarray) -> int:
    """Write the unit-8 stream: records of (8192 REAL*8, 8192 INT*4), the first zero
    label ends the stream.  Returns NREC, the number of records including the one
    holding the terminator (cval/poltwo.f:355-359)."""
    n = xx.size
    nrec = n // LSIZE + 1
    xb = np.zeros(nrec * LSIZE)
    ib = np.zeros(nrec * LSIZE, dtype=np.int32)
    xb[:n] = xx
    ib[:n] = ix
    with open(path, "wb") as f:
        for r in range(nrec):
            s = slice(r * LSIZE, (r + 1) * LSIZE)
            write_record(f, xb[s].astype("<f8").tobytes() + ib[s].astype("<i4").tobytes())
    return nrec


def read_pqrs(path, nrec: int | None = None):
    recs = read_records(path)
    if nrec is not None:
        recs = recs[:nrec]
    xx = np.concatenate([np.frombuffer(r[:8 * LSIZE], "<f8") for r in recs])
    ix = np.concatenate([np.frombuffer(r[8 * LSIZE:12 * LSIZE], "<i4") for r in recs])
    return xx, ix


# --------------------------------------------------------------------------- info file
def _i4(*v):
    return struct.pack("<%di" % len(v), *v)


def _r8(*v):
    return struct.pack("<%dd" % len(v), *v)


def _ch(s: str, n: int) -> bytes:
    return s.encode("ascii")[:n].ljust(n, b" ")


def write_info(path, n_ao: int, c_square: np.ndarray, h_packed: np.ndarray,
               s_packed: np.ndarray | None = None, enuc: float = 0.0, n_elec: int = 2,
               n_vec: int | None = None, group: str = "C1") -> None:
    """A C1 info file with N s-shells on one centre, as pshf writes it
    (pshf/daclos.f:41-59): record 1 = header + basis tables in the order rondof.f:98-106
    reads them; record 2 = S packed; record 3 = h packed; records 4-6 packed dummies;
    record 7 = C, N x N column-major."""
    n = n_ao
    nsh = n
    ngauss = n
    nat = 1
    nt = 1
    nvec = n if n_vec is None else n_vec
    head = _i4(*([0] * 19)) + _i4(nt, nsh, n, nvec, 0) + _ch("rhf", 8) \
        + _i4(ngauss, n_elec, 0, 1, n_elec // 2, n_elec // 2, nat) + _ch(group, 8) + _i4(2) \
        + _r8(*([0.0] * 10))
    assert len(head) == 56 * 4
    body = b"".join(_i4(1, k + 1, k + 1, 1, 1) for k in range(nsh))          # ktype,kstart,kloc,katom,kng
    body += b"".join(_r8(1.0 + k, 1.0, 0.0, 0.0, 0.0, 0.0) for k in range(ngauss))  # ex,cs,cp,cd,cf,cg
    body += _r8(0.0, 0.0, 0.0) + _r8(0.0) + _i4(1, 1, 1) + _r8(float(n_elec))   # c(3),a,ns,ks,iatno,zan
    body += _i4(*range(1, nsh + 1))                                           # newsh(i,1)
    body += _r8(*([1.0] * 9)) + _r8(*([1.0] * 36)) + _r8(*([1.0] * 100))       # ptr, dtr, ftr
    body += b"".join(_ch("X", 4) + _ch("%d" % (k + 1), 4) + _ch("s", 4) for k in range(n))  # iflab
    body += _i4(*([1] * (nsh * 12))) + _i4(*([0] * 48))                       # iso, invt
    body += _r8(enuc) + _ch("  pshf", 6)
    pa = n * (n + 1) // 2
    if s_packed is None:
        s_packed = np.eye(n)[np.tril_indices(n)]
    with open(path, "wb") as f:
        write_record(f, head + body)
        write_record(f, np.asarray(s_packed, "<f8").tobytes())
        write_record(f, np.asarray(h_packed, "<f8").tobytes())
        for _ in range(3):
            write_record(f, np.zeros(pa, "<f8").tobytes())
        write_record(f, np.asarray(c_square, dtype="<f8").reshape(-1, order="F").tobytes())


def write_info_molecule(path, mol, c_square: np.ndarray, h_packed: np.ndarray,
                        s_packed: np.ndarray | None = None, enuc: float = 0.0,
                        n_elec: int = 2, group: str | None = None) -> None:
    """An info file for a molecule with point-group symmetry (a synth_sym.Molecule):
    record 1 holds the header and the basis and symmetry tables in the order
    rondof.f:30-38 reads them -- shells (ktype, kstart, kloc, katom, kng), one primitive
    per shell, two atoms on the z axis, the image shell of every shell under each of the
    nt operations (1-based), the p / d / f transformation matrices (diagonal, 3x3, 6x6,
    10x10 per operation), AO labels, iso, invt, enuc, ante; records 2-7 as write_info."""
    n = mol.n_ao
    nsh = mol.nshell
    nt = mol.nt
    grp = group or {"C2v": "CNV", "D2h": "DNH", "Cinfv": "CINFV", "Dinfh": "DINFH"}[mol.group]
    head = _i4(*([0] * 19)) + _i4(nt, nsh, n, n, 0) + _ch("rhf", 8) \
        + _i4(nsh, n_elec, 0, 1, n_elec // 2, n_elec // 2, 2) + _ch(grp, 8) + _i4(2) \
        + _r8(*([0.0] * 10))
    assert len(head) == 56 * 4
    body = b"".join(_i4(kt, s + 1, mol.kloc[s] + 1, atom + 1, 1)
                    for s, (atom, kt) in enumerate(mol.shells))
    body += b"".join(_r8(1.0 + 0.5 * s, 1.0, 1.0, 1.0, 0.0, 0.0) for s in range(nsh))
    for atom, z in enumerate(mol.atom_z):
        body += _r8(0.0, 0.0, z) + _r8(0.0) + _i4(1, 1, 1) + _r8(float(n_elec) / 2)
    body += _i4(*[int(mol.newsh[s, g]) + 1 for g in range(nt) for s in range(nsh)])
    pd, dd = mol.ptr_diag()
    ptr = np.zeros((nt, 3, 3))
    dtr = np.zeros((nt, 6, 6))
    ftr = np.zeros((nt, 10, 10))
    for g in range(nt):
        ptr[g] = np.diag(pd[:, g])
        dtr[g] = np.diag(dd[:, g])
        ftr[g] = np.eye(10)
    # Fortran order: ptr(i, 3(g-1)+c) = M_g[i][c], i fastest
    body += _r8(*[ptr[g, i, c] for g in range(nt) for c in range(3) for i in range(3)])
    body += _r8(*[dtr[g, i, c] for g in range(nt) for c in range(6) for i in range(6)])
    body += _r8(*[ftr[g, i, c] for g in range(nt) for c in range(10) for i in range(10)])
    body += b"".join(_ch("AT%d" % (mol.shells[mol.ishell[a]][0] + 1), 4) + _ch("%d" % (a + 1), 4)
                     + _ch("spd"[mol.ktyp[a] - 1] + "%d" % mol.numc[a], 4) for a in range(n))
    body += _i4(*([1] * (nsh * 12))) + _i4(*([0] * 48))
    body += _r8(enuc) + _ch("  pshf", 6)
    pa = n * (n + 1) // 2
    if s_packed is None:
        s_packed = np.eye(n)[np.tril_indices(n)]
    with open(path, "wb") as f:
        write_record(f, head + body)
        write_record(f, np.asarray(s_packed, "<f8").tobytes())
        write_record(f, np.asarray(h_packed, "<f8").tobytes())
        for _ in range(3):
            write_record(f, np.zeros(pa, "<f8").tobytes())
        write_record(f, np.asarray(c_square, dtype="<f8").reshape(-1, order="F").tobytes())


# --------------------------------------------------------------------------- namelists
def ijfil_namelist(prefix: str = "") -> str:
    return "&ijfil prefix='%s' /\n" % prefix if prefix else "&ijfil /\n"


def ijklin_namelist(nrec: int, nvecf: int = 0, frozen: int = 0, ydp: bool = True,
                    precision: int | None = None, nslice: int | None = None,
                    istop: int | None = None, trec: float | None = None) -> str:
    """&ijklin for a C1 run (ijkl.f:28-33); precision / nslice are keys of the GPU host
    only (0 = FP64 MFMA, 1 = split FP16 with nslice slices)."""
    parts = ["ydp=%s" % ("T" if ydp else "F"), "nrec=%d" % nrec, "yauto=T", "chstan=F"]
    if nvecf:
        parts.append("nvecf=%d" % nvecf)
    if frozen:
        parts.append("ygel(1:%d)=%s" % (frozen, ",".join(["T"] * frozen)))
    if precision is not None:
        parts.append("precision=%d" % precision)
    if nslice is not None:
        parts.append("nslice=%d" % nslice)
    if istop is not None:
        parts.append("istop=%d" % istop)
    if trec is not None:
        parts.append("trec=%r" % trec)
    return "&ijklin " + ", ".join(parts) + " /\n"


# --------------------------------------------------------------------------- outputs
def read_f50(path, n_mo: int, n_vec_mono: int | None = None):
    """Split f50 into (mo_packed, last_record_tail, h_mo_packed).  The MO stream holds
    P_M(P_M+1)/2 valid doubles in 8192-double records, the last one written full
    length (tijkld.f:84-88); the final record is the one-electron block (tmono.f:56)."""
    recs = read_records(path)
    pm = n_mo * (n_mo + 1) // 2
    nval = pm * (pm + 1) // 2
    nrec = -(-nval // LSIZE)
    stream = np.concatenate([np.frombuffer(r, "<f8") for r in recs[:nrec]])
    mo = stream[:nval].copy()
    tail = stream[nval:].copy()
    hmo = np.frombuffer(recs[nrec], "<f8").copy() if len(recs) > nrec else None
    return dict(mo=mo, tail=tail, hmo=hmo, n_records=len(recs), n_mo_records=nrec,
                record_lengths=[len(r) for r in recs])


def write_f50(path, mo_packed: np.ndarray, hmo_packed: np.ndarray) -> int:
    """The f50 stream as tijkld/wtrbld and tmono write it: full 8192-double records; the
    last MO record is written full length, its tail holding what the previous record
    left in the buffer (tijkld.f:68-74,84-88); then one record with h_MO."""
    buf = np.zeros(LSIZE)
    n = mo_packed.size
    nrec = 0
    with open(path, "wb") as f:
        for r0 in range(0, n, LSIZE):
            cnt = min(LSIZE, n - r0)
            buf[:cnt] = mo_packed[r0:r0 + cnt]
            write_record(f, buf.astype("<f8").tobytes())
            nrec += 1
        write_record(f, np.asarray(hmo_packed, "<f8").tobytes())
    return nrec


def read_f25(path):
    recs = read_records(path)
    r0 = recs[0]
    nsym, nvec, noc, ntrans = struct.unpack_from("<4i", r0, 0)
    itsyv = list(struct.unpack_from("<%di" % nvec, r0, 16))
    off = 16 + 4 * nvec
    (norb,) = struct.unpack_from("<i", r0, off)
    ydp = r0[off + 4]
    (chstan,) = struct.unpack_from("<i", r0, off + 5)
    return dict(nsym=nsym, nvec=nvec, noc=noc, ntrans=ntrans, itsyv=itsyv, norb=norb,
                ydp=ydp, chstan=chstan, records=recs)
