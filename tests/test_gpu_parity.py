"""GPU parity of the HIP transform against the CPU oracle and the reference fixtures.

Bar: FP64 path within 1e-12 absolute of the oracle (which is bit-exact with the
reference program) on the same inputs; layout/ordering checks are exact.
"""
import os

import numpy as np
import pytest

import oracle
from cipsi_amd import formats, synth
from tests.conftest import load_golden

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


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def compacted_c(c_full, frozen):
    n = c_full.shape[0]
    if not frozen:
        return np.asfortranarray(c_full)
    return np.asfortranarray(np.concatenate([c_full[:, frozen:], c_full[:, n - frozen:]], axis=1))


@pytest.mark.parametrize("n,m", [(1, 1), (5, 3), (12, 12), (17, 9), (33, 33), (40, 24)])
def test_halves_match_tpqkld_tijkld(ctx, n, m):
    print(f"[halves] n={n} m={m}", flush=True)
    case = synth.make_case(n, m, max(2, 2 * n), 4.0, 100 + n + m)
    V, c = case["ao_packed"], case["c"]
    h_ref = oracle.tpqkld(V, c)
    h = ctx.tpqkld(dev(V), dev(c))
    torch.cuda.synchronize()
    assert np.max(np.abs(h.cpu().numpy() - h_ref)) <= TOL
    out_ref = oracle.tijkld(h_ref, c)
    out = ctx.tijkld(dev(h_ref), dev(c))
    torch.cuda.synchronize()
    assert np.max(np.abs(out.cpu().numpy() - out_ref)) <= TOL


@pytest.mark.parametrize("name", ["tiny_n12_frozen2", "sparse_n10_shuffled"])
def test_labels_path_matches_reference_fixture(ctx, name):
    g = load_golden(name)
    n = int(g["n_ao"])
    print(f"[labels] {name}", flush=True)
    case = synth.make_case(n, int(g["n_mo"]), int(g["rank"]), float(g["decay"]), int(g["seed"]),
                           frozen=int(g["frozen"]))
    xx, ix = formats.labels_from_packed(case["ao_packed"], n, cutoff=float(g["cutoff"]))
    if bool(g["shuffle"]):
        perm = np.random.default_rng(int(g["seed"]) + 3).permutation(xx.size)
        xx, ix = xx[perm], ix[perm]
    xx = np.append(xx, 0.0)
    ix = np.append(ix, 0).astype(np.int32)
    cc = compacted_c(g["c_full"], int(g["frozen"]))
    out = ctx.itijkl(xx, ix, cc)                       # host-pointer entry point
    assert np.max(np.abs(out - g["mo"])) <= TOL
    out_d = ctx.itijkl(dev(xx), dev(ix), dev(cc))       # device entry point
    torch.cuda.synchronize()
    assert np.max(np.abs(out_d.cpu().numpy() - g["mo"])) <= TOL
    # the unpacked block equals the oracle's transd exactly
    blk = ctx.transd(dev(xx), dev(ix), n)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(blk.cpu().numpy(), oracle.transd(n, xx, ix))
    np.testing.assert_allclose(ctx.tmono(g["h_packed"], cc), g["hmo"], rtol=0, atol=TOL)


def test_cfg0_n40_against_reference_and_oracle(ctx):
    g = load_golden("cfg0_n40")
    print("[cfg0] n=40", flush=True)
    case = synth.make_case(40, 40, 120, 8.0, 2024)
    c = compacted_c(g["c_full"], 0)
    out = ctx.transform(case["ao_packed"], c)
    ref = oracle.ijkl(case["ao_packed"], c)
    assert np.max(np.abs(out - ref)) <= TOL
    assert np.max(np.abs(out[g["sample_idx"]] - g["sample_val"])) <= TOL
    assert abs(np.sum(out) - g["full_sum"]) <= 1e-8


def test_cfg1_n120_frozen10_against_reference(ctx):
    g = load_golden("cfg1_n120_frozen10")
    print("[cfg1] n=120 m=110", flush=True)
    case = synth.make_case(120, 110, 360, 15.0, 2025, frozen=10)
    out = ctx.transform(dev(case["ao_packed"]), dev(case["c"]))
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    assert np.max(np.abs(o[g["sample_idx"]] - g["sample_val"])) <= TOL
    assert abs(np.sum(o) - g["active_sum"]) <= 1e-6
    assert abs(np.sum(o * o) - g["active_sumsq"]) <= 1e-6 * g["active_sumsq"]
    # whole stream against the independent closed form (different rounding path)
    cf = synth.mo_reference_closed_form(case["B"], case["c"])
    assert np.max(np.abs(o - cf)) <= 1e-11
    # whole stream against the oracle's whole transform (every AO pair through tpqkld,
    # every MO pair of the oracle's H through tijkld) at the 1e-12 bar
    nth = max(1, min(16, os.cpu_count() or 1))
    pa, pm = 120 * 121 // 2, 110 * 111 // 2
    h_ref = oracle.tpqkld_rows(case["ao_packed"], case["c"], np.arange(pa, dtype=np.int64),
                               nthreads=nth)
    o_ref = oracle.tijkld_rows(h_ref, case["c"], np.arange(pm, dtype=np.int64), nthreads=nth)
    d = float(np.max(np.abs(o - o_ref)))
    print(f"[cfg1] max|gpu - oracle| over all {o.size} integrals = {d:.3e}", flush=True)
    assert d <= TOL


def test_errors_reported_not_fatal(ctx):
    from cipsi_amd import ijkl
    c = np.eye(4)[:, :3]
    V = synth.make_case(4, 3, 4, 2.0, 1)["ao_packed"]
    V3 = synth.make_case(3, 3, 4, 2.0, 1)["ao_packed"]
    with pytest.raises(ijkl.IjklError) as e:
        ctx.transform(V3, np.eye(3, 4))            # n_mo > n_ao
    assert e.value.status == -1
    with pytest.raises(ijkl.IjklError) as e:
        bad = V.copy()
        bad[3] = np.nan
        ctx.transform(bad, c)
    assert e.value.status == -5
    with pytest.raises(ijkl.IjklError) as e:
        ctx.itijkl(np.array([1.0, 0.0]), np.array([formats.label_encode(5, 1, 1, 1), 0],
                                                   dtype=np.int32), c)
    assert e.value.status == -2
    # the context is still usable afterwards
    out = ctx.transform(V, c)
    assert np.max(np.abs(out - oracle.ijkl(V, c))) <= TOL


@pytest.mark.parametrize("n,m", [(40, 24), (97, 90), (200, 177)])
def test_every_tile_variant_gives_the_same_bits(ctx, n, m):
    """Every tile variant pair (waves, panel depth, occupancy hint) sums each element in
    the same order, so the packed output is bit-identical across plans; the tiles deal
    the column blocks of a matrix round robin, and ragged half-2 rows give tiles with
    every number of busy waves.  The smallest size is also checked against the oracle.
    A 12-wave variant is built only up to 12 row blocks (N_AO <= 192); above that the
    plan is refused with CIPSI_ERR_UNSUPPORTED."""
    from cipsi_amd import ijkl
    print(f"[variants] n={n} m={m}", flush=True)
    case = synth.make_case(n, m, 2 * n, 6.0, 300 + n)
    V, c = dev(case["ao_packed"]), dev(case["c"])
    ref = ctx.transform(V, c)
    torch.cuda.synchronize()
    if n <= 40:
        assert np.max(np.abs(ref.cpu().numpy() - oracle.ijkl(case["ao_packed"], case["c"]))) <= TOL
    shapes = ijkl.variants()
    nv = len(shapes)
    built = lambda v: shapes[v][0] != 12 or (n + 15) // 16 <= 12
    try:
        for v in range(nv):
            for v2 in (v, (v + 3) % nv):
                if not (built(v) and built(v2)):
                    with pytest.raises(ijkl.IjklError) as e:
                        ctx.set_plan(n, m, v, v2)
                    assert e.value.status == -7
                    continue
                ctx.set_plan(n, m, v, v2)
                out = ctx.transform(V, c)
                torch.cuda.synchronize()
                assert torch.equal(out, ref), f"variants ({v}, {v2})"
    finally:
        ctx.set_plan(n, m, *ctx.get_plan(1, 1))    # back to the context defaults


@pytest.mark.parametrize("n,m", [(8, 5), (60, 33), (144, 129), (250, 240), (256, 256)])
def test_paired_half2_gives_the_same_bits(ctx, n, m):
    """The default 8-wave tile runs a single transform's half 2 with the column blocks
    regrouped: a matrix of more than 8 blocks gives a full piece of 8 and a remainder, and
    two pieces of at most 8 blocks together share a work-group.  The sizes give rows of
    1 to 16 column blocks (remainders of every size, N_MO = 256 a remainder of 8).  The
    output equals the unpaired half 2 of the 8-wave, panel-depth-32 tile bit for bit, and
    the smallest size equals the oracle."""
    print(f"[paired half 2] n={n} m={m}", flush=True)
    case = synth.make_case(n, m, 2 * n, 8.0, 900 + n + m)
    V, c = dev(case["ao_packed"]), dev(case["c"])
    ref = ctx.transform(V, c)
    torch.cuda.synchronize()
    if n <= 8:
        assert np.max(np.abs(ref.cpu().numpy() - oracle.ijkl(case["ao_packed"], case["c"]))) <= TOL
    try:
        ctx.set_plan(n, m, 2, 3)                 # half 2 on the unpaired panel-depth-32 tile
        out = ctx.transform(V, c)
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
    finally:
        ctx.set_plan(n, m, *ctx.get_plan(1, 1))
    del V, out, ref
    torch.cuda.empty_cache()


def test_label_stream_duplicates_and_stale_tail(ctx):
    """transd semantics on the device, bit for bit against the oracle (transd.f:85-97,
    369-372): later duplicates overwrite earlier ones; the first zero label ends the
    stream, and what follows it (the stale tail of the last record: valid labels and
    labels outside the basis) is ignored; a label outside the basis before the end is an
    error.  The 4-byte-offset views take the unaligned (scalar) path of the claim pass."""
    from cipsi_amd import ijkl
    n = 9
    print("[labels-tail]", flush=True)
    case = synth.make_case(n, n, 18, 3.0, 410)
    xx, ix = formats.labels_from_packed(case["ao_packed"], n)
    rng = np.random.default_rng(411)
    perm = rng.permutation(xx.size)
    xx, ix = xx[perm], ix[perm]
    dup = rng.choice(xx.size, 40, replace=False)                  # later duplicates win
    xx = np.concatenate([xx, rng.standard_normal(40)])
    ix = np.concatenate([ix, ix[dup]])
    stale = rng.choice(xx.size, 30, replace=False)                # valid labels after the end
    tail_x = np.concatenate([rng.standard_normal(30), [1.0, 2.0]])
    tail_i = np.concatenate([ix[stale], formats.label_encode([n + 1, 3], [1, 0], [1, 1], [1, 1])])
    xx = np.concatenate([xx, [0.0], tail_x])
    ix = np.concatenate([ix, [0], tail_i]).astype(np.int32)
    want = oracle.transd(n, xx, ix)
    for off in (0, 1):
        x_d = dev(np.concatenate([[7.0] * off, xx]))[off:]
        i_d = dev(np.concatenate([np.array([5] * off, np.int32), ix]))[off:]
        got = ctx.transd(x_d, i_d, n)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got.cpu().numpy(), want)
    bad = ix.copy()
    bad[3] = formats.label_encode(n + 1, 1, 1, 1)
    with pytest.raises(ijkl.IjklError) as e:
        ctx.transd(dev(xx), dev(bad), n)
    assert e.value.status == -2
    got = ctx.transd(dev(xx), dev(ix), n)                        # still usable
    np.testing.assert_array_equal(got.cpu().numpy(), want)


def test_non_canonical_labels_stored_as_the_reference_does(ctx):
    """The reference's C1 route stores a label as read (transd.f:356,369-372): a label
    with i < j, k < l or ij < kl lands at pair0(pair0(i,j), pair0(k,l)) as given (with
    every field in 1..N_AO that position is always inside the block).  The device unpack
    equals the oracle's transd (the restatement pinned to the flang-built reference)
    bit for bit on such labels."""
    n = 7
    print("[labels-noncanonical]", flush=True)
    rng = np.random.default_rng(420)
    lab = [(2, 5, 1, 1), (4, 4, 3, 6), (1, 2, 7, 7), (3, 1, 2, 2), (6, 6, 6, 5), (1, 7, 2, 7)]
    xx = np.concatenate([rng.standard_normal(len(lab)), [0.0]])
    ix = np.array([formats.label_encode(*t) for t in lab] + [0], dtype=np.int32)
    want = oracle.transd(n, xx, ix)
    got = ctx.transd(dev(xx), dev(ix), n)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(got.cpu().numpy(), want)
    # (1,7|2,7) lands where (1,2|7,7) did and, being later, overwrites it
    assert np.count_nonzero(want) == len(lab) - 1


def test_transform_under_graph_capture(ctx):
    """A caller can capture the device transform in a CUDA graph (the library launches
    only kernels on the caller's stream once its workspace exists); each replay, with new
    inputs copied in and eager work in between, equals the eager result bit for bit."""
    n, m = 40, 24
    print(f"[graph] n={n} m={m}", flush=True)
    cases = [synth.make_case(n, m, 2 * n, 6.0, 700 + k) for k in range(3)]
    V, c = dev(cases[0]["ao_packed"]), dev(cases[0]["c"])
    refs = [ctx.transform(dev(k["ao_packed"]), dev(k["c"])) for k in cases]
    torch.cuda.synchronize()
    assert np.max(np.abs(refs[0].cpu().numpy() - oracle.ijkl(cases[0]["ao_packed"],
                                                             cases[0]["c"]))) <= TOL
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):                 # warm-up: workspace and tables exist
        ctx.transform(V, c)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        out = ctx.transform(V, c)
    for rep in range(2):
        for k, case in enumerate(cases):
            V.copy_(dev(case["ao_packed"]))
            c.copy_(dev(case["c"]))
            scratch = torch.randn(1 << 20, device="cuda", dtype=torch.float64)
            scratch.mul_(2.0)                  # eager work between replays
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(out, refs[k]), f"replay {rep} case {k}"


def test_captured_graph_survives_later_calls(ctx):
    """A captured transform keeps working after later eager calls on the same context
    that regrow its workspace (a larger N_AO), need other half-2 work tables (another
    N_MO, the split path's tile width) or change the size's tile plan: grown buffers are
    retired, not freed, and work tables are never rewritten, until release_workspace."""
    n, m = 40, 24
    print(f"[graph-later] n={n} m={m}", flush=True)
    case = synth.make_case(n, m, 2 * n, 6.0, 710)
    V, c = dev(case["ao_packed"]), dev(case["c"])
    ref = ctx.transform(V, c)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        ctx.transform(V, c)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        out = ctx.transform(V, c)
    g.replay()
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    plan = ctx.get_plan(n, m)
    try:
        big = synth.make_case(96, 80, 192, 6.0, 711)        # regrows vsq, h, cr
        ctx.transform(dev(big["ao_packed"]), dev(big["c"]))
        other_m = np.ascontiguousarray(case["c"][:, :30])   # a new (N_MO, width) table
        ctx.transform(V, dev(other_m))
        ctx.transform(V, c, precision="split_fp16")          # the split path's table
        ctx.set_plan(n, m, 0, 4)                             # a new plan for this size
        ctx.transform(V, c)
        torch.cuda.synchronize()
        for rep in range(2):
            out.zero_()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(out, ref), f"replay {rep} after later calls"
    finally:
        ctx.set_plan(n, m, *plan)


def test_host_entry_input_scan(ctx):
    """The host entry points scan the uploaded values on the device before computing: a
    NaN or an infinity anywhere (the odd last element too) is rejected with
    CIPSI_ERR_NONFINITE; in a label stream only the entries before the first zero label
    count (itijkl.f:62-68), so a NaN in the stale tail after it is accepted.  The
    split-FP16 error estimate scales with max |x|, so doubling the input doubles it."""
    from cipsi_amd import ijkl
    n, m = 6, 5                                  # 231 packed values: odd length
    case = synth.make_case(n, m, 2 * n, 3.0, 41)
    V, c = case["ao_packed"], case["c"]
    assert V.size % 2 == 1
    ref = oracle.ijkl(V, c)
    for pos, val in [(V.size - 1, np.inf), (0, np.nan), (V.size // 2, -np.inf)]:
        bad = V.copy()
        bad[pos] = val
        with pytest.raises(ijkl.IjklError) as e:
            ctx.transform(bad, c)
        assert e.value.status == -5, (pos, val)
    assert np.max(np.abs(ctx.transform(V, c) - ref)) <= TOL
    xx, ix = formats.labels_from_packed(V, n)
    tail_x = np.array([0.0, np.nan, 1.0])          # terminator, then a stale tail
    tail_i = np.array([0, formats.label_encode(2, 1, 1, 1), formats.label_encode(1, 1, 1, 1)],
                      dtype=np.int32)
    out = ctx.itijkl(np.concatenate([xx, tail_x]), np.concatenate([ix, tail_i]), c)
    assert np.max(np.abs(out - ref)) <= TOL
    bad_x = xx.copy()
    bad_x[-1] = np.nan                              # last entry before the terminator
    with pytest.raises(ijkl.IjklError) as e:
        ctx.itijkl(np.concatenate([bad_x, tail_x]), np.concatenate([ix, tail_i]), c)
    assert e.value.status == -5
    _, e1 = ctx.transform(V, c, precision="split_fp16", return_err=True)
    _, e2 = ctx.transform(2.0 * V, c, precision="split_fp16", return_err=True)
    assert e1 > 0 and e2 == pytest.approx(2.0 * e1, rel=1e-12)


@pytest.mark.parametrize("n,m", [(97, 90), (200, 177)])
def test_autotune_keeps_the_bits_and_built_variants(ctx, n, m):
    """cipsi_ijkl_autotune times the built variants of both halves on synthetic data of
    the size and stores the winners as the size's plan: the plan it returns is the one
    get_plan reads, it never names the 12-wave variant above 12 row blocks (N_AO 200),
    and the tuned transform gives the default plan's bits."""
    from cipsi_amd import ijkl
    print(f"[autotune] n={n} m={m}", flush=True)
    case = synth.make_case(n, m, 2 * n, 6.0, 900 + n)
    V, c = dev(case["ao_packed"]), dev(case["c"])
    ref = ctx.transform(V, c)
    torch.cuda.synchronize()
    shapes = ijkl.variants()
    try:
        v1, v2 = ctx.autotune(n, m)
        assert (v1, v2) == ctx.get_plan(n, m)
        for v in (v1, v2):
            assert 0 <= v < len(shapes)
            assert shapes[v][0] != 12 or (n + 15) // 16 <= 12
        out = ctx.transform(V, c)
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
    finally:
        ctx.set_plan(n, m, *ctx.get_plan(1, 1))


def test_empty_label_streams(ctx):
    """A stream with no entries, or one whose first label is the terminator, unpacks to
    an all-zero block (the block is zeroed first, itijkl.f:191-192, and the first zero
    label ends the stream, itijkl.f:62-68), so every MO integral is zero: host and device
    entry points."""
    n, m = 7, 5
    c = synth.make_case(n, m, 2 * n, 3.0, 51)["c"]
    nout = synth.n_pairs(synth.n_pairs(m))
    cases = [(np.zeros(0), np.zeros(0, dtype=np.int32)),
             (np.array([0.0, 2.5, 1.0]),
              np.array([0, formats.label_encode(1, 1, 1, 1), formats.label_encode(2, 1, 1, 1)],
                       dtype=np.int32))]
    for xx, ix in cases:
        out = ctx.itijkl(xx, ix, c)
        assert out.shape == (nout,) and not np.any(out)
        if xx.size:
            d = ctx.itijkl(dev(xx), torch.as_tensor(ix, device="cuda"), dev(c))
            torch.cuda.synchronize()
            assert d.numel() == nout and not torch.any(d)
