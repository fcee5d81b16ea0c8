"""The library's host-side MO preparation for molecules with symmetry
(cipsi_ijkl_sym_group, cipsi_ijkl_sym_prepare: otodon + rotat2 + vector) against the
reference's ijom bytes and MO irreps (tests/golden/sym_*.npz) and against the oracle's
restatement, plus the error tags rotat2 stops with.  Host code only: no GPU needed."""
import os
import struct

import numpy as np
import pytest

import oracle
from cipsi_amd import ijkl, synth_sym
from tests.conftest import load_golden
from tests.test_oracle_sym import SYM, f25_records, regenerate

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
