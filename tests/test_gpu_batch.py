"""Batched scan (cipsi_ijkl_batch_*): every geometry of a scan, transformed by the grouped
kernels in one graph launch, equals the single-transform path on the same
inputs bit for bit (the tile variants and row classes only change which work-group
computes an element, not the order of its sums), and the single path is pinned to the
oracle and the closed form by tests/test_gpu_parity.py and tests/test_gpu_fullsize.py.
Here the smallest geometry is also checked against the oracle directly, and one
geometry of the 64-geometry scan against the closed form.

Graph replays: the captured graph is replayed with new inputs and with eager torch
work and synchronisations between the replays; every replay is checked."""
import numpy as np
import pytest

import oracle
from cipsi_amd import synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from cipsi_amd import ijkl
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = ijkl.Context(0)
    yield c
    c.close()


def make_geoms(sizes, seed0, rank_factor=2, decay=6.0):
    aos, cs, Bs = [], [], []
    for i, (n, m) in enumerate(sizes):
        B = synth.eri_factors(n, rank_factor * n, decay, seed0 + i)
        Bs.append(B)
        aos.append(synth.eri_packed_torch(B, "cuda"))
        cs.append(torch.as_tensor(np.ascontiguousarray(synth.orthonormal(n, seed0 + 100 + i)[:, :m]),
                                  device="cuda"))
    return aos, cs, Bs


def outputs(sizes):
    return [torch.full((synth.n_pairs(synth.n_pairs(m)),), float("nan"), dtype=torch.float64,
                       device="cuda") for _, m in sizes]


def single(ctx, aos, cs):
    return [ctx.transform(a, c) for a, c in zip(aos, cs)]


MIXED = [(20, 12), (70, 62), (33, 33), (100, 92), (65, 57), (130, 122), (16, 5), (140, 140)]


def test_batch_matches_single_path_bitwise(ctx):
    from cipsi_amd import ijkl
    print(f"[batch-mixed] sizes={MIXED}", flush=True)
    aos, cs, _ = make_geoms(MIXED, 500)
    ref = single(ctx, aos, cs)
    plan = ijkl.BatchPlan(ctx, [n for n, _ in MIXED], [m for _, m in MIXED])
    classes = plan.classes()
    want = sorted({max(4, (n + 15) // 16) for n, _ in MIXED})   # 16-row blocks of N_AO
    assert [k[0] for k in classes] == want
    assert sum(k[1] for k in classes) == len(MIXED)
    for use_graph in (False, True):
        out = outputs(MIXED)
        plan.run(aos, cs, out, use_graph=use_graph)
        torch.cuda.synchronize()
        for g, (o, r) in enumerate(zip(out, ref)):
            assert torch.equal(o, r), f"geometry {g} {MIXED[g]} graph={use_graph}"
    # the smallest geometries against the oracle as well
    for g in (0, 6):
        n, m = MIXED[g]
        exp = oracle.ijkl(aos[g].cpu().numpy(), cs[g].cpu().numpy())
        assert np.max(np.abs(out[g].cpu().numpy() - exp)) <= 1e-12
    plan.close()


def test_paired_half2_in_the_batch(ctx):
    """Half 2 with the 8-wave, depth-16 variant runs the paired kernel over a whole row class:
    MO pairs of more than 8 column blocks split into a full piece and a remainder, and two
    pieces of one geometry share a work-group (N_MO = 256: remainders of 8 blocks).  Two
    geometries share row class 10.  Every plan with that half-2 variant gives the bits of the
    single transform, eager and from the graph; so does an unpaired half 2 afterwards."""
    from cipsi_amd import ijkl
    sizes = [(150, 144), (160, 139), (40, 33), (250, 240), (256, 256), (70, 62)]
    print(f"[batch-paired] sizes={sizes}", flush=True)
    aos, cs, _ = make_geoms(sizes, 950)
    ref = single(ctx, aos, cs)
    plan = ijkl.BatchPlan(ctx, [n for n, _ in sizes], [m for _, m in sizes])
    shapes = ijkl.variants()
    paired = shapes.index((8, 16, 1))
    unpaired = shapes.index((8, 32, 1))
    for v2 in (paired, unpaired):
        for k in range(len(plan.classes())):
            plan.set_class_plan(k, paired, v2)
        assert all(c[3] == v2 for c in plan.classes())
        for use_graph in (False, True):
            out = outputs(sizes)
            plan.run(aos, cs, out, use_graph=use_graph)
            torch.cuda.synchronize()
            for g, (o, r) in enumerate(zip(out, ref)):
                assert torch.equal(o, r), f"geometry {g} {sizes[g]} half2 {v2} graph={use_graph}"
        del out
    plan.close()


def test_batch_autotune_and_class_plans(ctx):
    from cipsi_amd import ijkl
    sizes = MIXED[:6]
    aos, cs, _ = make_geoms(sizes, 600)
    ref = single(ctx, aos, cs)
    plan = ijkl.BatchPlan(ctx, [n for n, _ in sizes], [m for _, m in sizes])
    out = outputs(sizes)
    classes = plan.autotune(aos, cs, out)
    torch.cuda.synchronize()
    nv = len(ijkl.variants())
    for rb, ng, v1, v2 in classes:
        assert 0 <= v1 < nv and 0 <= v2 < nv and ng >= 1
    for g in range(len(sizes)):
        assert torch.equal(out[g], ref[g]), f"autotune output, geometry {g}"
    # every explicit class plan gives the same bits
    for v in range(nv):
        for k in range(len(classes)):
            plan.set_class_plan(k, v, (v + 3) % nv)
        out = outputs(sizes)
        plan.run(aos, cs, out, use_graph=True)
        torch.cuda.synchronize()
        for g in range(len(sizes)):
            assert torch.equal(out[g], ref[g]), f"variant {v}, geometry {g}"
    plan.close()


def test_graph_replays_with_new_inputs(ctx):
    from cipsi_amd import ijkl
    sizes = [(24, 20), (72, 64), (40, 32)]
    plan = ijkl.BatchPlan(ctx, [n for n, _ in sizes], [m for _, m in sizes])
    for rep in range(4):
        print(f"[batch-replay] {rep}", flush=True)
        aos, cs, _ = make_geoms(sizes, 700 + 10 * rep)
        out = outputs(sizes)
        plan.run(aos, cs, out, use_graph=True)
        junk = torch.randn(1 << 20, device="cuda", dtype=torch.float64).sum()   # eager work
        torch.cuda.synchronize()
        ref = single(ctx, aos, cs)
        assert torch.isfinite(junk)
        for g in range(len(sizes)):
            assert torch.equal(out[g], ref[g]), f"replay {rep}, geometry {g}"
    plan.close()


def test_batch_rejects_bad_input(ctx):
    from cipsi_amd import ijkl
    with pytest.raises(ijkl.IjklError):
        ijkl.BatchPlan(ctx, [10, 20], [11, 5])          # n_mo > n_ao
    with pytest.raises(ijkl.IjklError):
        ijkl.BatchPlan(ctx, [300], [10])                # beyond the compiled row classes
    assert ijkl.batch_workspace_bytes([10], [12]) == -1
    plan = ijkl.BatchPlan(ctx, [10, 12], [8, 12])
    aos, cs, _ = make_geoms([(10, 8), (12, 12)], 800)
    with pytest.raises(ValueError):
        plan.run(aos[:1], cs[:1], outputs([(10, 8)]))
    plan.close()


def test_scan_64_geometries(ctx):
    """The 64-geometry scan (N_AO uniform in 60..180, N_MO = N_AO - 8, rank 3 N_AO):
    autotuned, one graph launch; sampled geometries equal the single path bit for bit
    and the smallest one matches the closed form (pq|rs) = sum_P Bt^P_pq Bt^P_rs."""
    from cipsi_amd import ijkl
    scan = synth.batch_scan()
    sizes = [(g["n_ao"], g["n_mo"]) for g in scan]
    print(f"[scan64] workspace {ijkl.batch_workspace_bytes(*zip(*sizes)) / 2**30:.1f} GiB",
          flush=True)
    aos, cs, Bs = [], [], []
    for g in scan:
        B = synth.eri_factors(g["n_ao"], g["rank"], g["decay"], g["seed"])
        aos.append(synth.eri_packed_torch(B, "cuda"))
        cs.append(torch.as_tensor(np.ascontiguousarray(
            synth.orthonormal(g["n_ao"], g["seed"] + 1)[:, :g["n_mo"]]), device="cuda"))
        Bs.append(B)
    plan = ijkl.BatchPlan(ctx, [n for n, _ in sizes], [m for _, m in sizes])
    out = outputs(sizes)
    print(f"[scan64] classes before tuning {plan.classes()}", flush=True)
    plan.autotune(aos, cs, out)
    print(f"[scan64] classes after tuning {plan.classes()}", flush=True)
    out = outputs(sizes)
    plan.run(aos, cs, out, use_graph=True)
    torch.cuda.synchronize()
    for o in out:
        assert bool(torch.isfinite(o).all())
    sample = sorted({0, 17, 42, 63, int(np.argmin([n for n, _ in sizes])),
                     int(np.argmax([n for n, _ in sizes]))})
    for g in sample:
        r = ctx.transform(aos[g], cs[g])
        assert torch.equal(out[g], r), f"geometry {g} {sizes[g]}"
    g = int(np.argmin([n for n, _ in sizes]))
    c = cs[g].cpu().numpy()
    cf = synth.mo_reference_closed_form(Bs[g], c)
    assert np.max(np.abs(out[g].cpu().numpy() - cf)) <= 1e-11
    plan.close()


def test_transform_batch_one_call(ctx):
    """cipsi_ijkl_transform_batch_dev, the one-call form: the first call for a size list
    builds and autotunes a plan on its data, a second call with the same sizes and new
    data replays the captured graph, and a new size list replaces the plan.  Every
    geometry equals the single-transform path bit for bit each time."""
    print("[batch one-call]", flush=True)
    for sizes, seed in ((MIXED[:5], 800), (MIXED[:5], 820), (MIXED[2:7], 840)):
        aos, cs, _ = make_geoms(sizes, seed)
        ref = single(ctx, aos, cs)
        out = ctx.transform_batch(aos, cs)
        torch.cuda.synchronize()
        for g in range(len(sizes)):
            assert torch.equal(out[g], ref[g]), f"seed {seed}, geometry {g}"
    ctx.release_workspace()                      # frees the kept plan too
    aos, cs, _ = make_geoms(MIXED[:2], 860)
    ref = single(ctx, aos, cs)
    out = ctx.transform_batch(aos, cs)
    torch.cuda.synchronize()
    for g in range(2):
        assert torch.equal(out[g], ref[g])


def test_batch_autotune_skips_variants_a_class_cannot_hold(ctx):
    """A plan with a geometry of 13 row blocks (N_AO = 208): the batch autotuner never
    picks the 12-wave variant for that class (it is built only up to 12 blocks), an
    explicit class plan naming it is refused with CIPSI_ERR_UNSUPPORTED, and the tuned
    outputs equal the single-transform path bit for bit."""
    from cipsi_amd import ijkl
    print("[batch 13 blocks]", flush=True)
    sizes = [(208, 200), (40, 30)]
    aos, cs, _ = make_geoms(sizes, 900)
    ref = single(ctx, aos, cs)
    plan = ijkl.BatchPlan(ctx, [n for n, _ in sizes], [m for _, m in sizes])
    out = outputs(sizes)
    shapes = ijkl.variants()
    wide = [v for v, s in enumerate(shapes) if s[0] == 12]
    classes = plan.autotune(aos, cs, out)
    torch.cuda.synchronize()
    for g in range(len(sizes)):
        assert torch.equal(out[g], ref[g]), f"geometry {g}"
    for k, (rb, ng, v1, v2) in enumerate(classes):
        if rb > 12:
            assert v1 not in wide and v2 not in wide
            for v in wide:
                with pytest.raises(ijkl.IjklError) as e:
                    plan.set_class_plan(k, v, v)
                assert e.value.status == -7
    plan.close()
