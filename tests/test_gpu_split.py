"""Split-FP16 (Ozaki) path: FP16 digit slices on FP16 MFMA with FP32 accumulation.

The FP64 path is the reference here (itself pinned to the oracle and the closed form).
The split path must reach <= 1e-10 max abs error on the packed (pq|rs) at the default
slice count (6) for every size up to N_AO = 240, and its error must fall by about the
digit base (256) per added slice, which pins the slice recombination: a wrong digit
weight or a dropped digit product shows up as an error at the level of the previous
slice count.  The energy E = 1/2 sum Gamma (pq|rs) with a trace-1 PSD pair density is
compared as well, and the host entry point's a-priori error estimate must bound the
measured error."""
import numpy as np
import pytest

from cipsi_amd import synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def ctx():
    from cipsi_amd import ijkl
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = ijkl.Context(0)
    yield c
    c.close()


def inputs(n, m, rank, decay, seed, frozen=0):
    B = synth.eri_factors(n, rank, decay, seed)
    ao = synth.eri_packed_torch(B, "cuda")
    c = synth.orthonormal(n, seed + 1)[:, frozen:frozen + m]
    return ao, torch.as_tensor(np.ascontiguousarray(c), device="cuda")


@pytest.mark.parametrize("n,m,rank,decay,seed,frozen", [
    (12, 9, 24, 3.0, 1, 0), (24, 20, 48, 4.0, 5, 0), (40, 40, 120, 8.0, 2024, 0),
    (65, 57, 130, 6.0, 7, 0), (120, 110, 360, 15.0, 2025, 10)])
def test_split_default_slices_within_tolerance(ctx, n, m, rank, decay, seed, frozen):
    print(f"[split] n={n} m={m}", flush=True)
    ao, c = inputs(n, m, rank, decay, seed, frozen)
    ref = ctx.transform(ao, c)
    out = ctx.transform(ao, c, precision="split_fp16")
    err = float((out - ref).abs().max())
    print(f"[split] n={n} m={m} maxabs={err:.3e} vmax={float(ref.abs().max()):.3e}", flush=True)
    assert err <= TOL


@pytest.mark.parametrize("n,m", [(256, 256), (256, 1), (33, 1), (130, 129), (17, 16), (200, 144)])
def test_split_edge_sizes(ctx, n, m):
    """The largest basis (N_AO = N_MO = 256: every 32-row tile and 128-column tile full,
    K = 256), a single MO, a basis one row past a 32-row tile, MO counts that end in the
    first rows of the second 128-column tile of the triangle (half-2 rows l >= 128), and a
    basis below one 32-row tile: <= 1e-10 against the FP64 path."""
    print(f"[split edge] n={n} m={m}", flush=True)
    ao, c = inputs(n, m, 2 * n, 10.0, 3000 + n + m)
    ref = ctx.transform(ao, c)
    out = ctx.transform(ao, c, precision="split_fp16")
    err = float((out - ref).abs().max())
    print(f"[split edge] n={n} m={m} maxabs={err:.3e} vmax={float(ref.abs().max()):.3e}",
          flush=True)
    assert err <= TOL
    assert float(ref.abs().max()) > 0.0


def test_split_error_falls_with_slices(ctx):
    ao, c = inputs(40, 36, 120, 8.0, 2024)
    ref = ctx.transform(ao, c)
    errs = []
    for s in (3, 4, 5, 6):
        out = ctx.transform(ao, c, precision="split_fp16", n_slices=s)
        errs.append(float((out - ref).abs().max()))
    print(f"[split] errors by slice count 3..6: {errs}", flush=True)
    for a, b in zip(errs, errs[1:]):
        assert b < a / 32          # about 2^-8 per slice; 32 leaves room for rounding
    assert errs[0] > 1e-7          # three slices are far from FP64: the slices are used


def test_split_n240_error_and_energy(ctx):
    """N_AO = N_MO = 240 (rank 720, decay 20): max abs error on (pq|rs) and on
    E = 1/2 sum_pqrs Gamma_pqrs (pq|rs) for a trace-1 PSD pair density."""
    n = m = 240
    print("[split] n=240", flush=True)
    ao, c = inputs(n, m, 720, 20.0, 2026)
    ref = ctx.transform(ao, c)
    out = ctx.transform(ao, c, precision="split_fp16")
    err = float((out - ref).abs().max())
    _, gam = synth.pair_density(m, 2029, rank=8)
    g = torch.as_tensor(gam, device="cuda")
    pm = synth.n_pairs(m)
    a, b = synth.pair_index_arrays(m)
    mult = torch.as_tensor(np.where(a == b, 1.0, 2.0), device="cuda")
    # 1/2 sum over all index orders = 1/2 sum_{A,B} m(A) m(B) Gamma_AB g_AB over the
    # full pair supermatrix; the packed stream holds A >= B once, so weight 2 off the
    # diagonal of the supermatrix
    ar = torch.arange(pm, device="cuda")
    rows = torch.repeat_interleave(ar, ar + 1)
    cols = torch.arange(synth.n_pairs(pm), device="cuda") - rows * (rows + 1) // 2
    w = torch.where(rows == cols, 1.0, 2.0).to(torch.float64) * mult[rows] * mult[cols] * g
    e_ref = 0.5 * float(torch.sum(w * ref))
    e_split = 0.5 * float(torch.sum(w * out))
    print(f"[split] n=240 maxabs={err:.3e} E={e_ref:.15e} dE={abs(e_split - e_ref):.3e}",
          flush=True)
    assert err <= TOL
    assert abs(e_split - e_ref) <= TOL


@pytest.mark.parametrize("n,m,rank,decay,seed", [
    (40, 40, 120, 8.0, 2024), (121, 111, 363, 15.0, 2025), (240, 240, 720, 20.0, 2026)])
def test_split_full_mantissa_integrals(ctx, n, m, rank, decay, seed):
    """AO integrals with full 53-bit mantissas (unrounded ERI factors).  The grid
    integrals of the other tests are multiples of 2^-34 below 2^13, which six base-256
    digits of each AO row hold exactly; these carry bits below every fixed grid, so the
    digit truncation of the AO operand (half 1, step 1) is exercised.  Still <= 1e-10
    against the FP64 path, and the FP64 path still agrees with the oracle on a sample."""
    import oracle
    print(f"[split full-mantissa] n={n} m={m}", flush=True)
    B = synth.eri_factors(n, rank, decay, seed, grid=False)
    ao = synth.eri_packed_torch(B, "cuda")
    scaled = ao * 2.0**34
    assert bool((scaled != torch.round(scaled)).any()), "integrals are on the 2^-34 grid"
    del scaled
    c = torch.as_tensor(np.ascontiguousarray(synth.orthonormal(n, seed + 1)[:, :m]), device="cuda")
    ref = ctx.transform(ao, c)
    out = ctx.transform(ao, c, precision="split_fp16")
    err = float((out - ref).abs().max())
    print(f"[split full-mantissa] n={n} m={m} maxabs={err:.3e} "
          f"vmax={float(ref.abs().max()):.3e}", flush=True)
    assert err <= TOL
    if n <= 40:
        o_ref = oracle.ijkl(ao.cpu().numpy(), c.cpu().numpy())
        assert float(np.max(np.abs(ref.cpu().numpy() - o_ref))) <= 1e-12


def test_split_host_entry_error_estimate(ctx):
    ao, c = inputs(30, 26, 60, 5.0, 11)
    ao_h, c_h = ao.cpu().numpy(), c.cpu().numpy()
    ref = ctx.transform(ao_h, c_h)
    out, est = ctx.transform(ao_h, c_h, precision="split_fp16", return_err=True)
    err = float(np.max(np.abs(out - ref)))
    print(f"[split] host n=30: maxabs={err:.3e} estimate={est:.3e}", flush=True)
    assert err <= est <= 1e-6


def test_split_under_graph_capture(ctx):
    """The split device transform launches only kernels on the caller's stream once its
    workspace and triangle tables exist, so a caller can capture it: each replay, with new
    inputs copied in and eager work in between, equals the eager split result bit for
    bit (both halves, step 1 and step 2, are deterministic)."""
    n, m = 72, 66
    print(f"[split graph] n={n} m={m}", flush=True)
    cases = [inputs(n, m, 2 * n, 6.0, 720 + k) for k in range(3)]
    V = cases[0][0].clone()
    c = cases[0][1].clone()
    refs = [ctx.transform(a, b, precision="split_fp16") for a, b in cases]
    fp64 = ctx.transform(cases[0][0], cases[0][1])
    torch.cuda.synchronize()
    assert float((refs[0] - fp64).abs().max()) <= TOL
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):                 # warm-up: workspace and tables exist
        ctx.transform(V, c, precision="split_fp16")
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        out = ctx.transform(V, c, precision="split_fp16")
    for rep in range(2):
        for k, (a, b) in enumerate(cases):
            V.copy_(a)
            c.copy_(b)
            scratch = torch.randn(1 << 20, device="cuda", dtype=torch.float64)
            scratch.mul_(2.0)                  # eager work between replays
            out.fill_(float("nan"))
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(out, refs[k]), f"replay {rep} case {k}"


def test_split_rejects_bad_slice_counts(ctx):
    from cipsi_amd import ijkl
    ao, c = inputs(12, 10, 24, 3.0, 3)
    for s in (1, 2, 7):
        with pytest.raises(ijkl.IjklError):
            ctx.transform(ao, c, precision="split_fp16", n_slices=s)
