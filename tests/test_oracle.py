"""Pins the CPU oracle (oracle/ijkl_oracle.c) before it is trusted as the checker.

* against the reference ijkl program itself (golden fixtures made by
  tests/golden/make_golden.py from a flang build of CIPSI_sources/ijkl): bit-exact;
* against mathematics that fixes the answer independently of the loop nests:
  the closed form through the ERI factors, C = identity, C = signed permutation,
  truncated identity, a dense brute-force einsum, norm preservation;
* planted mistakes (a transposed C, a dropped diagonal doubling) must fail them.
"""
import hashlib

import numpy as np
import pytest

import oracle
from cipsi_amd import formats, synth
from tests.conftest import load_golden


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def compacted_c(c_full, frozen):
    """Columns the reference transforms for NVECF = N with YGEL(1:frozen): active
    columns compacted to the front, the last `frozen` columns left in place
    (rotat2.f:50-60; nvec = nvecf, ijkl.f:187)."""
    n = c_full.shape[0]
    if not frozen:
        return np.asfortranarray(c_full)
    return np.asfortranarray(np.concatenate([c_full[:, frozen:], c_full[:, n - frozen:]], axis=1))


def golden_inputs(g):
    n = int(g["n_ao"])
    case = synth.make_case(n, int(g["n_mo"]), int(g["rank"]), float(g["decay"]), int(g["seed"]),
                           frozen=int(g["frozen"]))
    V = case["ao_packed"]
    assert sha(V) == str(g["ao_sha"]), "synthetic AO integrals drifted from the fixture"
    return case, V


@pytest.mark.parametrize("name", ["tiny_n12_frozen2", "sparse_n10_shuffled"])
def test_oracle_bitexact_small_full_stream(name):
    g = load_golden(name)
    n = int(g["n_ao"])
    case, V = golden_inputs(g)
    xx, ix = formats.labels_from_packed(V, n, cutoff=float(g["cutoff"]))
    if bool(g["shuffle"]):
        perm = np.random.default_rng(int(g["seed"]) + 3).permutation(xx.size)
        xx, ix = xx[perm], ix[perm]
    assert sha(ix) == str(g["labels_sha"])
    block = oracle.transd(n, np.append(xx, 0.0), np.append(ix, 0).astype(np.int32))
    if float(g["cutoff"]) == 0.0:
        np.testing.assert_array_equal(block, V)
    cc = compacted_c(g["c_full"], int(g["frozen"]))
    np.testing.assert_array_equal(np.asarray(g["ijom_c"]).reshape(n, n, order="F"), cc)
    out = oracle.ijkl(block, cc)
    np.testing.assert_array_equal(out, g["mo"])
    np.testing.assert_array_equal(oracle.tmono(g["h_packed"], cc), g["hmo"])


@pytest.mark.parametrize("name", ["cfg0_n40"])
def test_oracle_bitexact_cfg0(name):
    g = load_golden(name)
    case, V = golden_inputs(g)
    out = oracle.ijkl(V, compacted_c(g["c_full"], int(g["frozen"])))
    np.testing.assert_array_equal(out[g["sample_idx"]], g["sample_val"])
    assert np.sum(out) == g["full_sum"]
    assert np.sum(out * out) == g["full_sumsq"]
    nrec = int(g["n_mo_records"])
    np.testing.assert_array_equal(out[(nrec - 1) * formats.LSIZE:], g["last_record"])


@pytest.mark.slow
def test_oracle_bitexact_cfg1_frozen():
    """N_AO=120, 10 frozen: the reference transforms 110 compacted + 10 stale columns; the
    leading P_110(P_110+1)/2 outputs are the 110-MO transform (ilk-outer ordering)."""
    g = load_golden("cfg1_n120_frozen10")
    case, V = golden_inputs(g)
    n, m, fr = int(g["n_ao"]), int(g["n_mo"]), int(g["frozen"])
    out_active = oracle.ijkl(V, case["c"])                       # 110 active columns only
    np.testing.assert_array_equal(out_active[g["sample_idx"]], g["sample_val"])
    assert np.sum(out_active) == g["active_sum"]
    assert np.sum(out_active * out_active) == g["active_sumsq"]


def test_stale_tail_and_records():
    g = load_golden("tiny_n12_frozen2")
    n = int(g["n_ao"])
    pm = n * (n + 1) // 2
    nval = pm * (pm + 1) // 2
    assert int(g["n_mo_records"]) == -(-nval // formats.LSIZE)
    assert int(g["n_records"]) == int(g["n_mo_records"]) + 1
    assert g["tail"].size == int(g["n_mo_records"]) * formats.LSIZE - nval


# ---------------------------------------------------------------- mathematics pins
def small_case(n=9, m=7, rank=14, seed=5):
    case = synth.make_case(n, m, rank, 3.0, seed)
    return case


def test_closed_form_matches_loops():
    case = small_case()
    out = oracle.ijkl(case["ao_packed"], case["c"])
    ref = synth.mo_reference_closed_form(case["B"], case["c"])
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-12)


def dense_eri(V, n):
    pa = n * (n + 1) // 2
    full = synth.unpack_sym(V, pa)
    a, b = synth.pair_index_arrays(n)
    g = np.zeros((n, n, n, n))
    pair = np.zeros((n, n), dtype=np.int64)
    pair[a, b] = np.arange(pa)
    pair[b, a] = np.arange(pa)
    g[:, :, :, :] = full[pair[:, :, None, None], pair[None, None, :, :]]
    return g


def pack_mo(gmo, m):
    pm = m * (m + 1) // 2
    a, b = synth.pair_index_arrays(m)
    sup = gmo[a[:, None], b[:, None], a[None, :], b[None, :]]      # [ilk-ish][nij]
    return sup[np.tril_indices(pm)]


def test_bruteforce_einsum():
    n, m = 6, 5
    case = synth.make_case(n, m, 10, 2.0, 9)
    g = dense_eri(case["ao_packed"], n)
    c = case["c"]
    gmo = np.einsum("pqrs,pi,qj,rk,sl->ijkl", g, c, c, c, c)
    np.testing.assert_allclose(oracle.ijkl(case["ao_packed"], c), pack_mo(gmo, m), atol=1e-12)


def test_identity_and_permutation():
    n = 7
    case = synth.make_case(n, n, 12, 3.0, 21)
    V = case["ao_packed"]
    np.testing.assert_array_equal(oracle.ijkl(V, np.eye(n)), V)
    perm = np.random.default_rng(1).permutation(n)
    signs = np.where(np.random.default_rng(2).random(n) < 0.5, -1.0, 1.0)
    P = np.zeros((n, n))
    P[perm, np.arange(n)] = signs
    g = dense_eri(V, n)
    gp = np.einsum("pqrs,pi,qj,rk,sl->ijkl", g, P, P, P, P)
    np.testing.assert_array_equal(oracle.ijkl(V, P), pack_mo(gp, n))
    # truncated identity: the leading sub-block
    m = 4
    out = oracle.ijkl(V, np.eye(n)[:, :m])
    pm = m * (m + 1) // 2
    np.testing.assert_array_equal(out, V[:pm * (pm + 1) // 2])


def test_norm_preserved_square_orthonormal():
    case = small_case(n=8, m=8)
    V = case["ao_packed"]
    out = oracle.ijkl(V, case["c"])
    n = 8
    g = dense_eri(V, n)
    gm = dense_eri(out, n)
    assert abs(np.sum(g * g) - np.sum(gm * gm)) < 1e-9 * np.sum(g * g)


def test_planted_mistakes_fail():
    """A transposed C or a dropped diagonal doubling in the unpack must be caught."""
    case = small_case(n=6, m=6, seed=8)
    V, c = case["ao_packed"], case["c"]
    good = synth.mo_reference_closed_form(case["B"], c)
    assert np.max(np.abs(oracle.ijkl(V, np.asfortranarray(c.T)) - good)) > 1e-3
    xx, ix = formats.labels_from_packed(V, 6)
    i, j, k, l = formats.label_decode(ix)
    xx_bad = xx.copy()
    xx_bad[i == j] *= 0.5                  # pretend the producer halved twice
    blk = oracle.transd(6, np.append(xx_bad, 0.0), np.append(ix, 0).astype(np.int32))
    assert np.max(np.abs(oracle.ijkl(blk, c) - good)) > 1e-3


def test_transd_terminator_and_duplicates():
    n = 4
    V = synth.make_case(n, n, 6, 2.0, 3)["ao_packed"]
    xx, ix = formats.labels_from_packed(V, n)
    # terminator in the middle: everything after it is ignored
    cut = xx.size // 2
    blk = oracle.transd(n, np.concatenate([xx[:cut], [0.0], xx[cut:]]),
                        np.concatenate([ix[:cut], [0], ix[cut:]]).astype(np.int32))
    np.testing.assert_array_equal(blk[:cut], V[:cut])
    assert np.all(blk[cut:] == 0.0)
    # a later duplicate overwrites
    xx2 = np.append(xx, [123.0, 0.0])
    ix2 = np.append(ix, [ix[5], 0]).astype(np.int32)
    blk = oracle.transd(n, xx2, ix2)
    i, j, k, l = formats.label_decode(ix[5:6])
    f = 2.0 ** ((i == j).astype(int) + (k == l).astype(int) + ((i == k) & (j == l)).astype(int))
    assert blk[5] == 123.0 * f[0]
    # out-of-range label is rejected
    with pytest.raises(ValueError):
        oracle.transd(n, np.array([1.0, 0.0]), np.array([formats.label_encode(5, 1, 1, 1), 0],
                                                         dtype=np.int32))
