"""Point-group symmetry on the GPU (cipsi_ijkl_sym_transform_labels) against the
reference fixtures and the symmetric oracle: C2v, D2h, and the linear groups C-inf-v and
D-inf-h with their symmetry transformations.

The COMMON state the entry point takes (classes, AO subsets, packed C) comes from the
oracle's restatement of otodon / rotat2 / vector, which reproduces the reference's ijom
bytes (tests/test_oracle_sym.py).  Bar: every value of the f50 stream within 1e-12 of
the reference's / the oracle's, the block directory (ipair, nad, ktypsy) identical.
"""
import numpy as np
import pytest

import oracle
from cipsi_amd import synth_sym
from tests.conftest import load_golden
from tests.test_oracle_sym import LIN, SYM, regenerate

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    from cipsi_amd import ijkl
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = ijkl.Context(0)
    yield c
    c.close()


def prepared(case, c_square, h_packed, noc):
    t = synth_sym.tables(case["mol"])
    n = case["n_ao"]
    prep = oracle.sym_prepare(t, np.eye(n)[np.tril_indices(n)], h_packed, c_square, noc)
    gpu = dict(n_ao=n, nsym=prep.nsym, nt=t["nt"], its=prep.its, isymp=prep.isymp, nbp=prep.nbp,
               nbo=prep.nbo, ind=prep.ind, ivecsy=prep.ivecsy, c=prep.c_packed,
               ishell=t["ishell"], kloc=t["kloc"], ktyp=t["ktyp"], numc=t["numc"],
               newsh=t["newsh"], iptr=t["iptr"], idtr=t["idtr"], iftr=t["iftr"])
    ao = dict(ishell=t["ishell"], kloc=t["kloc"], ktyp=t["ktyp"], numc=t["numc"],
              newsh=t["newsh"], iptr=t["iptr"], idtr=t["idtr"], iftr=t["iftr"])
    return prep, gpu, ao


def with_end(xx, ix):
    return np.append(xx, 0.0), np.append(ix, 0).astype(np.int32)


@pytest.mark.parametrize("name", SYM)
def test_sym_gpu_matches_reference_fixture(ctx, name):
    print(f"[sym] {name}", flush=True)
    g = load_golden(name)
    case = regenerate(g)
    prep, tabs, ao = prepared(case, g["c_square"], g["h_packed"], int(g["noc"]))
    xx, ix = with_end(case["xx"], case["ix"])
    out, ipair, nad, ktypsy = ctx.sym_itijkl(tabs, xx, ix)
    o_ref, ip_ref, nad_ref, kt_ref = oracle.sym_itijkl(prep, ao, xx, ix)
    assert out.size == o_ref.size
    assert np.array_equal(ipair, ip_ref)
    assert np.array_equal(nad, nad_ref[:nad.size]) and np.array_equal(ktypsy, kt_ref[:ktypsy.size])
    ref = g["f50_stream"][:out.size]
    d_ref = float(np.max(np.abs(out - ref)))
    d_or = float(np.max(np.abs(out - o_ref)))
    print(f"[sym] {name}: {out.size} integrals in {ipair.shape[0]} blocks, "
          f"max|gpu - reference| {d_ref:.2e}, max|gpu - oracle| {d_or:.2e}", flush=True)
    assert d_ref <= TOL and d_or <= TOL


@pytest.mark.parametrize("grp,sa,sb,rank,seed,noc", [
    ("D2h", [1, 1, 1, 2, 2, 2, 3, 3, 3], None, 2, 41, 9),
    ("C2v", [1, 1, 1, 2, 2, 2, 3, 3, 3, 3], [1, 1, 2, 2, 3, 3], 2, 43, 11),
])
def test_sym_gpu_larger_against_oracle(ctx, grp, sa, sb, rank, seed, noc):
    """Bases of 50-54 AOs (irreps up to 24 AOs: two row blocks of the tiles) against the
    symmetric oracle on the same prepared state."""
    case = synth_sym.make_sym_case(grp, sa, sb, rank_per_irrep=rank, seed=seed)
    print(f"[sym-large] {grp} N={case['n_ao']}", flush=True)
    prep, tabs, ao = prepared(case, case["c_square"], case["h_packed"], noc)
    xx, ix = with_end(case["xx"], case["ix"])
    out, ipair, nad, ktypsy = ctx.sym_itijkl(tabs, xx, ix)
    o_ref, ip_ref, _, _ = oracle.sym_itijkl(prep, ao, xx, ix)
    assert np.array_equal(ipair, ip_ref)
    d = float(np.max(np.abs(out - o_ref)))
    print(f"[sym-large] {grp} N={case['n_ao']}: {out.size} integrals, max|gpu - oracle| {d:.2e}",
          flush=True)
    assert out.size == o_ref.size and d <= TOL


def test_sym_label_order_duplicates_stale_tail_and_empty_class_quartet(ctx):
    """transd's stream semantics with images: a shuffled stream, later duplicates (with
    other values) overwriting earlier ones, a stale tail after the terminator, and one
    class quartet whose integrals are all left out of the stream -- its blocks are then
    skipped (no directory entry, no output) -- all as the oracle does."""
    g = load_golden("sym_d2h_n20")
    case = regenerate(g)
    prep, tabs, ao = prepared(case, g["c_square"], g["h_packed"], int(g["noc"]))
    rng = np.random.default_rng(77)
    xx, ix = case["xx"].copy(), case["ix"].copy()
    # drop every label whose four AOs lie in the class of the first irrep only
    from cipsi_amd import formats
    i, j, k, l = formats.label_decode(ix)
    member = prep.ivecsy[0] != 0                      # AOs of class 1 (Ag / B1u here)
    only1 = member[i - 1] & member[j - 1] & member[k - 1] & member[l - 1]
    keep = ~only1
    xx, ix = xx[keep], ix[keep]
    perm = rng.permutation(xx.size)
    xx, ix = xx[perm], ix[perm]
    dup = rng.choice(xx.size, 25, replace=False)
    xx = np.concatenate([xx, rng.standard_normal(25)])
    ix = np.concatenate([ix, ix[dup]])
    xx = np.concatenate([xx, [0.0], rng.standard_normal(10)])
    ix = np.concatenate([ix, [0], ix[:10]]).astype(np.int32)
    out, ipair, nad, ktypsy = ctx.sym_itijkl(tabs, xx, ix)
    o_ref, ip_ref, nad_ref, kt_ref = oracle.sym_itijkl(prep, ao, xx, ix)
    full = oracle.sym_itijkl(prep, ao, *with_end(case["xx"], case["ix"]))[1]
    assert ip_ref.shape[0] < full.shape[0]            # a class quartet was skipped
    assert np.array_equal(ipair, ip_ref)
    assert np.array_equal(nad, nad_ref[:nad.size]) and np.array_equal(ktypsy, kt_ref[:ktypsy.size])
    assert out.size == o_ref.size
    assert float(np.max(np.abs(out - o_ref))) <= TOL


def test_sym_rejects_bad_input(ctx):
    from cipsi_amd import ijkl
    g = load_golden("sym_c2v_n19")
    case = regenerate(g)
    prep, tabs, ao = prepared(case, g["c_square"], g["h_packed"], int(g["noc"]))
    xx, ix = with_end(case["xx"], case["ix"])
    bad = ix.copy()
    bad[5] = formats_label(30, 1, 1, 1)
    with pytest.raises(ijkl.IjklError) as e:
        ctx.sym_itijkl(tabs, xx, bad)
    assert e.value.status == -2
    t2 = dict(tabs)
    t2["isymp"] = np.asarray(tabs["isymp"])[::-1].copy()      # not non-decreasing
    with pytest.raises(ijkl.IjklError) as e:
        ctx.sym_itijkl(t2, xx, ix)
    assert e.value.status == -1
    out, _, _, _ = ctx.sym_itijkl(tabs, xx, ix)               # still usable
    assert float(np.max(np.abs(out - g["f50_stream"][:out.size]))) <= TOL


def formats_label(i, j, k, l):
    from cipsi_amd import formats
    return formats.label_encode(i, j, k, l)


# ----------------------------------------------------------------------- linear groups
def prepared_linear(case, c_square, h_packed, noc, nvecf=None):
    t = synth_sym.tables(case["mol"])
    n = case["n_ao"]
    nsym, its, _, isydeg, so = oracle.sym_linear(synth_sym.GROUP_NAME[case["mol"].group], t)
    prep = oracle.sym_prepare_so(n, nsym, its, so, np.eye(n)[np.tril_indices(n)], h_packed,
                                 c_square, noc, nvecf=nvecf)
    gpu = dict(n_ao=n, nsym=prep.nsym, nt=t["nt"], its=prep.its, isymp=prep.isymp, nbp=prep.nbp,
               nbo=prep.nbo, ind=prep.ind, ivecsy=prep.ivecsy, c=prep.c_packed,
               ishell=t["ishell"], kloc=t["kloc"], ktyp=t["ktyp"], numc=t["numc"],
               newsh=t["newsh"], iptr=t["iptr"], idtr=t["idtr"], iftr=t["iftr"], isydeg=isydeg)
    ao = dict(ishell=t["ishell"], kloc=t["kloc"], ktyp=t["ktyp"], numc=t["numc"],
              newsh=t["newsh"], iptr=t["iptr"], idtr=t["idtr"], iftr=t["iftr"])
    return prep, gpu, ao, isydeg


@pytest.mark.parametrize("name", LIN)
def test_lin_gpu_matches_reference_fixture(ctx, name):
    """C-inf-v / D-inf-h: the blocks the symmetry transformations leave out are not in
    the stream, the directory (ipair, the signed nad pointers, ktypsy) is the reference's,
    and every integral is within 1e-12 of the reference's f50 stream."""
    print(f"[lin] {name}", flush=True)
    g = load_golden(name)
    case = regenerate(g)
    prep, tabs, ao, isydeg = prepared_linear(case, g["c_square"], g["h_packed"], int(g["noc"]))
    xx, ix = with_end(case["xx"], case["ix"])
    out, ipair, nad, ktypsy = ctx.sym_itijkl(tabs, xx, ix)
    o_ref, ip_ref, nad_ref, kt_ref = oracle.sym_itijkl(prep, ao, xx, ix, isydeg=isydeg)
    assert out.size == o_ref.size
    assert np.array_equal(ipair, ip_ref)
    assert np.array_equal(nad, nad_ref[:nad.size]) and np.array_equal(ktypsy, kt_ref[:ktypsy.size])
    ref = g["f50_stream"][:out.size]
    d_ref = float(np.max(np.abs(out - ref)))
    d_or = float(np.max(np.abs(out - o_ref)))
    print(f"[lin] {name}: {out.size} integrals in {ipair.shape[0]} blocks, "
          f"max|gpu - reference| {d_ref:.2e}, max|gpu - oracle| {d_or:.2e}", flush=True)
    assert d_ref <= TOL and d_or <= TOL


@pytest.mark.parametrize("grp,sa,sb,seed,noc", [
    ("Dinfh", [1, 1, 1, 2, 2, 2, 3, 3, 3], None, 51, 9),
    ("Cinfv", [1, 1, 1, 2, 2, 2, 3, 3, 3, 3], [1, 1, 2, 2, 3, 3], 53, 11),
])
def test_lin_gpu_larger_against_oracle(ctx, grp, sa, sb, seed, noc):
    """Linear bases of 50-54 AOs against the oracle on the same prepared state."""
    case = synth_sym.make_linear_case(grp, sa, sb, rank=2, seed=seed)
    print(f"[lin-large] {grp} N={case['n_ao']}", flush=True)
    prep, tabs, ao, isydeg = prepared_linear(case, case["c_square"], case["h_packed"], noc)
    xx, ix = with_end(case["xx"], case["ix"])
    out, ipair, nad, ktypsy = ctx.sym_itijkl(tabs, xx, ix)
    o_ref, ip_ref, nad_ref, kt_ref = oracle.sym_itijkl(prep, ao, xx, ix, isydeg=isydeg)
    assert np.array_equal(ipair, ip_ref)
    assert np.array_equal(nad, nad_ref[:nad.size]) and np.array_equal(ktypsy, kt_ref[:ktypsy.size])
    d = float(np.max(np.abs(out - o_ref)))
    print(f"[lin-large] {grp} N={case['n_ao']}: {out.size} integrals, max|gpu - oracle| {d:.2e}",
          flush=True)
    assert out.size == o_ref.size and d <= TOL


@pytest.mark.parametrize("y_first", [False, True])
def test_lin_gpu_unequal_partner_irreps(ctx, y_first):
    """nvecf cutting a pi pair in two leaves pi_x and pi_y with different MO counts; a
    block the rotation maps onto itself with its pairs exchanged is still stored as a
    triangle, whose rows stop at the shorter side's pair count (tijkld.f:75-77).  With
    the y partner first in record 7 the longer side is the other one."""
    case = synth_sym.make_linear_case("Cinfv", [1, 2, 2, 3], [1, 2], rank=2, seed=61)
    c = np.array(case["c_square"])
    irr = np.array(case["mo_irrep"])
    n = case["n_ao"]
    pairs = [k for k in range(n - 1) if irr[k] == 1 and irr[k + 1] == 2]
    if y_first:
        for k in pairs:
            c[:, [k, k + 1]] = c[:, [k + 1, k]]
    cut = pairs[len(pairs) // 2] + 1                     # keep the first member only
    print(f"[lin-cut] N={n} nvecf={cut} y_first={y_first}", flush=True)
    prep, tabs, ao, isydeg = prepared_linear(case, c, case["h_packed"], 2, nvecf=cut)
    assert prep.nbo[1] != prep.nbo[2]
    xx, ix = with_end(case["xx"], case["ix"])
    out, ipair, nad, ktypsy = ctx.sym_itijkl(tabs, xx, ix)
    o_ref, ip_ref, nad_ref, kt_ref = oracle.sym_itijkl(prep, ao, xx, ix, isydeg=isydeg)
    assert np.array_equal(ipair, ip_ref)
    assert np.array_equal(nad, nad_ref[:nad.size]) and np.array_equal(ktypsy, kt_ref[:ktypsy.size])
    assert out.size == o_ref.size
    assert float(np.max(np.abs(out - o_ref))) <= TOL
