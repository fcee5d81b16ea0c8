"""The library's host-side MO preparation for molecules with symmetry
(cipsi_ijkl_sym_group, cipsi_ijkl_sym_prepare: otodon + rotat2 + vector; for the linear
groups cipsi_ijkl_sym_linear and cipsi_ijkl_sym_prepare_so) against the reference's ijom
and f25 bytes (tests/golden/sym_*.npz, lin_*.npz) and against the oracle's restatement,
plus the error tags rotat2 stops with.  Host code only: no GPU needed."""
import os
import struct

import numpy as np
import pytest

import oracle
from cipsi_amd import ijkl, synth_sym
from tests.conftest import load_golden
from tests.test_oracle_sym import LIN, SYM, f25_records, regenerate

pytestmark = pytest.mark.skipif(not os.path.exists(ijkl.LIB_PATH), reason="library not built")


def test_group_tables_match_the_synthetic_groups():
    for grp, name, n in (("C2v", "CNV", 4), ("D2h", "DNH", 8)):
        nsym, its, jsym = ijkl.sym_group(name, 2, n)
        assert nsym == n
        assert np.array_equal(its, synth_sym.its_table(n))
        assert np.array_equal(jsym, np.array(synth_sym._CHI[grp]))
    assert ijkl.sym_group("cs", 0, 2)[0] == 2 and ijkl.sym_group("C1", 1, 1)[0] == 1
    for name, naxis, status in (("CINFV", 2, -7), ("DINFH", 2, -7), ("CNV", 3, -7), ("XYZ", 2, -1)):
        with pytest.raises(ijkl.IjklError) as e:
            ijkl.sym_group(name, naxis, 4)
        assert e.value.status == status


@pytest.mark.parametrize("name", SYM)
def test_prepare_matches_reference_and_oracle(name):
    g = load_golden(name)
    case = regenerate(g)
    t = synth_sym.tables(case["mol"])
    n = case["n_ao"]
    p = ijkl.sym_prepare(t, None, g["h_packed"], g["c_square"], int(g["noc"]))
    assert np.array_equal(p["c"].ravel(order="F"), g["ijom_c"])           # reference bytes
    rec0 = f25_records(g)[0]
    nvec = struct.unpack_from("<i", rec0, 4)[0]
    assert np.array_equal(np.frombuffer(rec0[16:16 + 4 * nvec], "<i4"), p["itsyv"])
    o = oracle.sym_prepare(t, np.eye(n)[np.tril_indices(n)], g["h_packed"], g["c_square"],
                           int(g["noc"]))
    assert np.array_equal(o.c, p["c"]) and np.array_equal(o.c_packed, p["c_packed"])
    for k in ("isymp", "nbp", "nbo", "ind"):
        assert np.array_equal(getattr(o, k), p[k]), k
    assert np.array_equal(o.ivecsy, p["ivecsy"]) and o.nsymp == p["nsymp"]
    # S given explicitly (the identity here) gives the same bits as yatom's unit S
    p2 = ijkl.sym_prepare(t, np.eye(n)[np.tril_indices(n)], g["h_packed"], g["c_square"],
                          int(g["noc"]))
    assert np.array_equal(p2["c"], p["c"])


def test_prepare_error_tags():
    """rotat2's stops: an MO mixing two irreps overlaps too many symmetry vectors (tag 1),
    two equal MOs are not orthogonal (tag 1), a zero column overlaps no irrep (tag 2), a
    column of norm 1e-7 still classified with trec = 1e-9 fails the norm check (tag -1),
    and a degenerate pair of one irrep overlaps fewer irreps than its size (tag 2).  The
    oracle stops with the same tags."""
    g = load_golden("sym_c2v_n19")
    case = regenerate(g)
    t = synth_sym.tables(case["mol"])
    n = case["n_ao"]
    c0 = np.asfortranarray(g["c_square"])
    irr = case["mo_irrep"]
    a = int(np.flatnonzero(irr == 0)[0])
    b = int(np.flatnonzero(irr != 0)[0])
    mixed = c0.copy()
    mixed[:, a] = (c0[:, a] + c0[:, b]) / np.sqrt(2.0)
    dup = c0.copy()
    dup[:, b] = c0[:, a]
    zero = c0.copy()
    zero[:, a] = 0.0
    tiny = c0.copy()
    tiny[:, a] = 1e-7 * c0[:, a]
    # three consecutive MOs of one energy, two of one irrep and one of another: the level
    # overlaps two irreps, fewer than its three members and more than one
    lvl = np.concatenate([np.flatnonzero(irr == irr[a])[:2], [b]])
    h = case["energies"].copy()
    order = np.concatenate([lvl, [k for k in range(n) if k not in lvl]])
    e = h[order].copy()
    e[1] = e[2] = e[0]
    deg_c = c0[:, order]
    deg_h = (deg_c @ np.diag(e) @ deg_c.T)[np.tril_indices(n)]
    # and two of one irrep alone: an accidental degeneracy, accepted without rotation
    acc_c = c0[:, order]
    e2 = h[order].copy()
    e2[1] = e2[0]
    acc_h = (acc_c @ np.diag(e2) @ acc_c.T)[np.tril_indices(n)]
    p_acc = ijkl.sym_prepare(t, None, acc_h, acc_c, int(g["noc"]))
    o_acc = oracle.sym_prepare(t, np.eye(n)[np.tril_indices(n)], acc_h, acc_c, int(g["noc"]))
    assert np.array_equal(p_acc["c"], o_acc.c)
    cases = [(mixed, g["h_packed"], 1e-5, 1), (dup, g["h_packed"], 1e-5, 1),
             (zero, g["h_packed"], 1e-5, 2), (tiny, g["h_packed"], 1e-9, -1),
             (deg_c, deg_h, 1e-5, 2)]
    for c, hp, trec, tag in cases:
        with pytest.raises(ijkl.IjklError) as e:
            ijkl.sym_prepare(t, None, hp, c, int(g["noc"]), trec=trec)
        assert e.value.status == -6 and e.value.tag == tag, (tag, e.value.tag)
        with pytest.raises(ValueError) as e2:
            oracle.sym_prepare(t, np.eye(n)[np.tril_indices(n)], hp, c, int(g["noc"]), trec=trec)
        assert e2.value.args[0] == tag


# ----------------------------------------------------------------------- linear groups
@pytest.mark.parametrize("name", LIN)
def test_linear_prepare_matches_reference_and_oracle(name):
    """otodon's linear-group tables (its and the isydeg records of f25), rotat2's
    coefficients (ijom) and the MO irreps, equal to the reference's bytes; every table
    equal to the oracle's."""
    g = load_golden(name)
    case = regenerate(g)
    t = synth_sym.tables(case["mol"])
    n = case["n_ao"]
    grp = synth_sym.GROUP_NAME[str(g["group"])]
    nsym, its, ntrans, isydeg, so = ijkl.sym_linear(grp, t)
    o = oracle.sym_linear(grp, t)
    assert (nsym, ntrans) == (o[0], o[2])
    assert np.array_equal(its, o[1]) and np.array_equal(isydeg, o[3]) and np.array_equal(so, o[4])
    recs = f25_records(g)
    assert struct.unpack_from("<i", recs[0], 0)[0] == nsym
    assert struct.unpack_from("<i", recs[0], 12)[0] == ntrans
    assert np.array_equal(np.frombuffer(recs[1], "<i4").reshape(nsym, nsym).T, its)
    for k in range(ntrans):
        assert np.array_equal(np.frombuffer(recs[2 + k], "<i4")[1:], isydeg[:, k])
    p = ijkl.sym_prepare_so(n, nsym, so, None, g["h_packed"], g["c_square"], int(g["noc"]))
    assert np.array_equal(p["c"].ravel(order="F"), g["ijom_c"])
    nvec = struct.unpack_from("<i", recs[0], 4)[0]
    assert np.array_equal(np.frombuffer(recs[0][16:16 + 4 * nvec], "<i4"), p["itsyv"])
    op = oracle.sym_prepare_so(n, nsym, its, so, np.eye(n)[np.tril_indices(n)], g["h_packed"],
                               g["c_square"], int(g["noc"]))
    assert np.array_equal(op.c_packed, p["c_packed"]) and op.nsymp == p["nsymp"]
    for k in ("isymp", "nbp", "nbo", "ind", "ivecsy"):
        assert np.array_equal(getattr(op, k), p[k]), k


def _linear_basis(group, shells_a, shells_b=None, z=(1.25, -1.25)):
    """Basis tables of a linear molecule with any shell types (f included): shells of
    atom 1 then atom 2 (D-inf-h: the same shells, sigma_h exchanging the atoms)."""
    if group == "DINFH":
        shells_b = list(shells_a)
    shells = [(0, k) for k in shells_a] + [(1, k) for k in (shells_b or [])]
    kloc, pos = [], 1
    for _, kt in shells:
        kloc.append(pos)
        pos += 2 * kt - 1
    na = len(shells_a)
    newsh = np.array([[s + 1, (s + na if at == 0 else s - na) + 1 if group == "DINFH" else s + 1]
                      for s, (at, _) in enumerate(shells)], dtype=np.int32)
    return dict(norb=pos - 1, ktype=np.array([k for _, k in shells], dtype=np.int32),
                kloc=np.array(kloc, dtype=np.int32), zsh=np.array([z[at] for at, _ in shells]),
                newsh=newsh)


@pytest.mark.parametrize("group,sa,sb,nsym,ntrans", [
    ("CINFV", [1, 2, 3, 4], [2, 1], 7, 2), ("DINFH", [1, 2, 4], None, 14, 2),
    ("CINFV", [1, 3], [1], 5, 2), ("DINFH", [3], None, 10, 2), ("DINFH", [1, 2], None, 10, 1),
    ("CINFV", [1, 1], [1], 5, 0)])
def test_linear_tables_with_f_shells_match_oracle(group, sa, sb, nsym, ntrans):
    """The product tables its and the transformations isydeg come from the classes'
    parities and the rotations; the oracle holds the reference's DATA tables literally.
    With f shells (nsym 7 and 14, no reference fixture) and with d but no p shell (otodon
    counts a p transformation whenever d shells are present)."""
    t = _linear_basis(group, sa, sb)
    got = ijkl.sym_linear(group, t)
    ref = oracle.sym_linear(group, t)
    assert got[0] == ref[0] == nsym and got[2] == ref[2] == ntrans
    for a, b in zip(got[1:], ref[1:]):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def test_linear_centre_atom():
    """D-inf-h with an atom at the centre (z = 0): its AOs enter only the classes of their
    own parity, with no partner; the library and the oracle agree."""
    t = _linear_basis("DINFH", [1, 2, 3], None, z=(0.0, 0.0))
    t["newsh"][:, 1] = np.arange(1, t["ktype"].size + 1)
    got = ijkl.sym_linear("DINFH", t)
    ref = oracle.sym_linear("DINFH", t)
    assert np.array_equal(got[4], ref[4])
    so = got[4]
    assert np.all((so != 0).sum(axis=0) == 1)       # every AO in one class
    assert np.all(so >= 0)


def test_linear_bad_arguments():
    t = _linear_basis("DINFH", [1, 2], None)
    with pytest.raises(ijkl.IjklError) as e:
        ijkl.sym_linear("DINFV", t)
    assert e.value.status == -1
    bad = dict(t, newsh=t["newsh"][:, :1])              # D-inf-h needs the sigma_h images
    with pytest.raises(ijkl.IjklError):
        ijkl.sym_linear("DINFH", bad)
    bad = dict(t, kloc=t["kloc"] + 3)                    # a shell past the last AO
    with pytest.raises(ijkl.IjklError):
        ijkl.sym_linear("DINFH", bad)
