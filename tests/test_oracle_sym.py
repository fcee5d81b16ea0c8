"""The symmetric oracle (oracle/ijkl_sym_oracle.c) against the reference and against
mathematics, on molecules with point-group symmetry (C2v heteronuclear and D2h
homonuclear diatomics, cipsi_amd.synth_sym).

Pins:
* bit for bit against the flang-built reference's own outputs (tests/golden/sym_*.npz):
  the MO coefficients rotat2 leaves (ijom), the MO irreps (f25 header), the whole f50
  stream of every symmetry block, the block directory (ipair, nad, ktypsy) and the
  h_MO record;
* against the closed form (pq|rs)_MO = sum_P Bt^P_pq Bt^P_rs, Bt^P = C^T B^P C, laid out
  block by block in the order the reference's directory lists the blocks
  (itijkl.f:208-213) and, inside a block, l, k outer and j, i inner (tpqkld.f:83-102,
  tijkld.f:31-79): an independent route to every value of the stream;
* planted mistakes fail: a flipped p-component sign in the operation tables changes the
  symmetry images transd generates, and the closed-form check catches it.
"""
import struct

import numpy as np
import pytest

import oracle
from cipsi_amd import synth, synth_sym
from tests.conftest import load_golden

SYM = ["sym_c2v_n19", "sym_c2v_deg_n19", "sym_d2h_n20", "sym_d2h_n24", "sym_c2v_n30",
       "sym_d2h_n38"]


def regenerate(g):
    import hashlib
    sb = list(g["shells_b"]) if g["shells_b"].size else None
    case = synth_sym.make_sym_case(str(g["group"]), list(g["shells_a"]), sb,
                                   rank_per_irrep=int(g["rank"]), seed=int(g["seed"]),
                                   degenerate_pair=bool(g["degenerate"]))
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    assert sha(case["ao_packed"]) == str(g["ao_sha"])
    assert sha(case["xx"]) == str(g["xx_sha"]) and sha(case["ix"]) == str(g["labels_sha"])
    return case


def ao_tables(t):
    return dict(ishell=t["ishell"], kloc=t["kloc"], ktyp=t["ktyp"], numc=t["numc"],
                newsh=t["newsh"], iptr=t["iptr"], idtr=t["idtr"], iftr=t["iftr"])


def run_oracle(case, g, tables=None):
    t = tables or synth_sym.tables(case["mol"])
    prep = oracle.sym_prepare(t, np.eye(case["n_ao"])[np.tril_indices(case["n_ao"])],
                              g["h_packed"], g["c_square"], int(g["noc"]))
    xx = np.append(case["xx"], 0.0)
    ix = np.append(case["ix"], 0).astype(np.int32)
    return prep, oracle.sym_itijkl(prep, ao_tables(t), xx, ix)


def f25_records(g):
    from cipsi_amd import formats
    import tempfile, os
    with tempfile.NamedTemporaryFile(delete=False) as f:
        f.write(g["f25_bytes"].tobytes())
        path = f.name
    try:
        return formats.read_records(path)
    finally:
        os.unlink(path)


def closed_form_stream(case, prep, ipair):
    """The f50 stream from the dense MO integrals, block by block in directory order."""
    n = case["n_ao"]
    ia, ib = np.tril_indices(n)
    bt = []
    for P in range(case["B"].shape[0]):
        b = np.zeros((n, n))
        b[ia, ib] = case["B"][P]
        b[ib, ia] = case["B"][P]
        bt.append(prep.c.T @ b @ prep.c)
    bt = np.array(bt)
    g4 = np.einsum("Ppq,Prs->pqrs", bt, bt)
    mos = [np.flatnonzero(prep.itsyv == s + 1) for s in range(prep.nsym)]
    out = []
    for nij, nkl in ipair:
        ii = int((1 + np.sqrt(8 * nij - 7)) // 2)
        while ii * (ii - 1) // 2 >= nij:
            ii -= 1
        jj = nij - ii * (ii - 1) // 2
        kk = int((1 + np.sqrt(8 * nkl - 7)) // 2)
        while kk * (kk - 1) // 2 >= nkl:
            kk -= 1
        ll = nkl - kk * (kk - 1) // 2
        A, B, C, D = mos[ii - 1], mos[jj - 1], mos[kk - 1], mos[ll - 1]
        diag = ii == kk and jj == ll
        ilk = 0
        for l in range(A.size):
            for k in range(l + 1 if ii == jj else B.size):
                ilk += 1
                nij_ = 0
                stop = False
                for j in range(C.size):
                    for i in range(j + 1 if kk == ll else D.size):
                        nij_ += 1
                        out.append(g4[A[l], B[k], C[j], D[i]])
                        if diag and nij_ == ilk:
                            stop = True
                            break
                    if stop:
                        break
    return np.array(out)


@pytest.mark.parametrize("name", SYM)
def test_sym_oracle_bitexact_against_reference(name):
    g = load_golden(name)
    case = regenerate(g)
    prep, (stream, ipair, nad, ktypsy) = run_oracle(case, g)
    # rotat2 / vector: coefficients and irreps
    assert np.array_equal(prep.c.ravel(order="F"), g["ijom_c"])
    recs = f25_records(g)
    nsym, nvec, noc, ntrans = struct.unpack_from("<4i", recs[0], 0)
    assert nsym == prep.nsym and nvec == prep.nvec and ntrans == 0
    assert np.array_equal(np.frombuffer(recs[0][16:16 + 4 * nvec], "<i4"), prep.itsyv)
    assert np.array_equal(np.frombuffer(recs[1], "<i4").reshape(nsym, nsym).T, prep.its)
    # the MO stream of every block, then the directory
    ref = g["f50_stream"]
    assert stream.size <= ref.size < stream.size + 8192
    assert np.array_equal(stream, ref[:stream.size])
    npaire = struct.unpack_from("<i", recs[2], 0)[0]
    assert npaire == ipair.shape[0]
    assert np.array_equal(np.frombuffer(recs[2][4:], "<i4").reshape(-1, 2), ipair)
    nsyt = struct.unpack_from("<i", recs[3], 0)[0]
    assert np.array_equal(np.frombuffer(recs[3][4:], "<i2"), nad[:nsyt])
    assert np.array_equal(np.frombuffer(recs[4], "<i2"), ktypsy[:nsyt])
    assert np.array_equal(oracle.tmono(g["h_packed"], prep.c), g["hmo"])


@pytest.mark.parametrize("name", SYM)
def test_sym_stream_matches_closed_form(name):
    g = load_golden(name)
    case = regenerate(g)
    prep, (stream, ipair, _, _) = run_oracle(case, g)
    cf = closed_form_stream(case, prep, ipair)
    assert cf.size == stream.size
    assert np.max(np.abs(cf - stream)) <= 1e-12


def test_sym_classes_and_blocks():
    """D2h homonuclear: the AOs of Ag and B1u (and of the three other pairs) coincide, so
    rotat2 finds four classes and makes each pair of irreps contiguous; C2v: four
    classes of one irrep each.  Only blocks whose irrep product is totally symmetric
    (its) appear in the directory."""
    for name, nsymp in (("sym_d2h_n20", 4), ("sym_c2v_n19", 4)):
        g = load_golden(name)
        case = regenerate(g)
        prep, (_, ipair, _, _) = run_oracle(case, g)
        assert prep.nsymp == nsymp
        assert np.all(np.diff(prep.isymp) >= 0)
        for nij, nkl in ipair:
            ii = max(a for a in range(1, prep.nsym + 1) if a * (a - 1) // 2 < nij)
            jj = nij - ii * (ii - 1) // 2
            kk = max(a for a in range(1, prep.nsym + 1) if a * (a - 1) // 2 < nkl)
            ll = nkl - kk * (kk - 1) // 2
            its = prep.its
            assert its[its[ii - 1, jj - 1] - 1, its[kk - 1, ll - 1] - 1] == 1


def test_planted_image_sign_fails():
    """Flip the sign of the x component of p functions under one atom-exchanging D2h
    operation: the images transd generates for px integrals change sign, and the stream
    no longer matches the closed form."""
    g = load_golden("sym_d2h_n20")
    case = regenerate(g)
    t = synth_sym.tables(case["mol"])
    t["iptr"] = t["iptr"].copy()
    t["iptr"][0, 2] = -t["iptr"][0, 2]
    prep, (stream, ipair, _, _) = run_oracle(case, g, tables=t)
    cf = closed_form_stream(case, prep, ipair)
    assert cf.size == stream.size
    assert np.max(np.abs(cf - stream)) > 1e-6
