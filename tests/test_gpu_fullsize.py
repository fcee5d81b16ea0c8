"""Full-size checks (N_AO = N_MO = 240, rank 720) and stream ordering with torch.

Parity at the headline size is pinned by
* the whole packed MO stream (418,197,660 integrals) against the oracle's whole
  transform -- tpqkld over every AO pair, then tijkld over every MO pair of the
  oracle's own H, the same restatement that is bit-exact with the reference program,
  run on the host cores with OpenMP -- at the 1e-12 bar;
* the whole stream against the closed form (pq|rs) = sum_P Bt^P_pq Bt^P_rs,
  Bt^P = C^T B^P C, computed on the GPU with torch (independent rounding path);
* every row of the GPU's H against the oracle's H.
"""
import os

import numpy as np
import pytest

import oracle
from cipsi_amd import synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

# host threads for the oracle's whole transforms (the GPU box gives one GPU 16 cores)
ORACLE_THREADS = max(1, min(16, os.cpu_count() or 1))


def oracle_whole(ao_host, c):
    """The oracle's whole transform, in parallel over AO pairs and then MO pairs; every
    element is computed with the serial restatement's own order of operations."""
    n, m = c.shape
    pa, pm = n * (n + 1) // 2, m * (m + 1) // 2
    h = oracle.tpqkld_rows(ao_host, c, np.arange(pa, dtype=np.int64), nthreads=ORACLE_THREADS)
    out = np.empty(pm * (pm + 1) // 2)
    oracle.tijkld_rows(h, c, np.arange(pm, dtype=np.int64), nthreads=ORACLE_THREADS, out=out)
    return h, out


@pytest.fixture(scope="module")
def ctx():
    from cipsi_amd import ijkl
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = ijkl.Context(0)
    yield c
    c.close()


def closed_form_packed(B, c):
    """Packed MO stream from the ERI factors, on the GPU."""
    n, m = c.shape
    a, b = np.tril_indices(n)
    am, bm = np.tril_indices(m)
    Bd = torch.as_tensor(B, device="cuda")
    cd = torch.as_tensor(np.ascontiguousarray(c), device="cuda")
    ia = torch.as_tensor(a, device="cuda")
    ib = torch.as_tensor(b, device="cuda")
    out_f = torch.empty((B.shape[0], am.size), dtype=torch.float64, device="cuda")
    for p0 in range(0, B.shape[0], 64):
        p1 = min(B.shape[0], p0 + 64)
        full = torch.zeros((p1 - p0, n, n), dtype=torch.float64, device="cuda")
        full[:, ia, ib] = Bd[p0:p1]
        full[:, ib, ia] = Bd[p0:p1]
        t = torch.matmul(torch.matmul(cd.t().contiguous(), full), cd)
        out_f[p0:p1] = t[:, torch.as_tensor(am, device="cuda"), torch.as_tensor(bm, device="cuda")]
    return synth.eri_packed_torch(out_f.cpu().numpy(), "cuda")


def test_stream_ordering_without_device_sync(ctx):
    """Results read through torch's own stream (.cpu()) without a device-wide sync."""
    n = 160
    print(f"[stream] n={n}", flush=True)
    B = synth.eri_factors(n, 2 * n, 10.0, 31)
    V = synth.eri_packed_torch(B, "cuda")
    c = synth.orthonormal(n, 32)
    h = ctx.tpqkld(V, torch.as_tensor(c, device="cuda")).cpu().numpy()
    a, b = np.tril_indices(n)
    for pq in (0, 1234, n * (n + 1) // 2 - 1):
        row = B[:, pq] @ B
        Z = np.zeros((n, n))
        Z[a, b] = row
        Z[b, a] = row
        O = c.T @ Z @ c
        assert np.max(np.abs(h[pq] - O[a, b])) <= 1e-11


def test_cfg2_n240_fp64_full_stream(ctx):
    n = m = 240
    print("[cfg2] n=240 fp64 full stream", flush=True)
    cfg = synth.CONFIGS[2]
    B = synth.eri_factors(n, cfg["rank"], cfg["decay"], cfg["seed"])
    V = synth.eri_packed_torch(B, "cuda")
    c = synth.orthonormal(n, cfg["seed"] + 1)
    cd = torch.as_tensor(c, device="cuda")
    out = ctx.transform(V, cd)
    ref = closed_form_packed(B, c)
    diff = float(torch.max(torch.abs(out - ref)))
    print(f"  max|gpu - closed form| = {diff:.3e}, max|out| = {float(torch.max(torch.abs(ref))):.3f}",
          flush=True)
    assert diff <= 1e-11
    del ref
    # the whole stream and every row of H against the oracle's whole transform
    h_gpu = ctx.tpqkld(V, cd).cpu().numpy()
    V_host = V.cpu().numpy()
    o = out.cpu().numpy()
    del V, out
    print(f"  oracle whole transform on {ORACLE_THREADS} host threads", flush=True)
    h_ref, o_ref = oracle_whole(V_host, c)
    dh = float(np.max(np.abs(h_ref - h_gpu)))
    del h_gpu, h_ref
    do = float(np.max(np.abs(o_ref - o)))
    print(f"  max|gpu - oracle|: H {dh:.3e}, packed MO stream {do:.3e} "
          f"({o.size} integrals)", flush=True)
    assert o.size == 418197660
    assert dh <= 1e-12
    assert do <= 1e-12


@pytest.mark.parametrize("n,m", [(72, 64), (90, 90), (100, 92), (136, 128), (170, 160),
                                 (200, 190), (216, 216), (256, 256)])
def test_every_row_class_against_closed_form(ctx, n, m):
    """One size per compiled row-block count not covered above (5, 6, 7, 9, 11, 13, 14 and
    the largest, 16 = N_AO 256): the whole packed stream against the closed form, and the
    first and last MO pairs of the second half against the oracle's tijkld."""
    print(f"[row class] n={n} m={m} blocks={(n + 15) // 16}", flush=True)
    B = synth.eri_factors(n, 2 * n, 12.0, 4000 + n)
    V = synth.eri_packed_torch(B, "cuda")
    c = np.ascontiguousarray(synth.orthonormal(n, 4100 + n)[:, :m])
    cd = torch.as_tensor(c, device="cuda")
    out = ctx.transform(V, cd)
    ref = closed_form_packed(B, c)
    diff = float(torch.max(torch.abs(out - ref)))
    print(f"  max|gpu - closed form| = {diff:.3e}", flush=True)
    assert diff <= 1e-11
    del ref
    h = ctx.tpqkld(V, cd).cpu().numpy()
    pm = m * (m + 1) // 2
    kl_list = np.array([0, pm // 2, pm - 1], dtype=np.int64)
    o_ref = oracle.tijkld_rows(h, c, kl_list)
    o = out.cpu().numpy()
    idx = np.concatenate([np.arange(k * (k + 1) // 2, k * (k + 1) // 2 + k + 1) for k in kl_list])
    assert np.max(np.abs(o_ref[idx] - o[idx])) <= 1e-12


def test_host_entry_points_over_many_bounce_chunks(ctx):
    """At N_AO = N_MO = 140 the packed blocks are 390 MB, seven 64 MiB chunks of the
    host entry points' pinned bounce slots (four slots, so each is reused), and the
    label stream is 390 MB of values plus 195 MB of labels.  The host entry points
    (packed and label stream) give the device path's bits -- the packed one with its
    expand run behind the upload and its output copied during half 2, into a fresh and
    into a reused output array -- and the label stream is the pshf stream of the same
    block."""
    from cipsi_amd import formats
    n = m = 140
    print(f"[bounce] n={n}", flush=True)
    B = synth.eri_factors(n, 2 * n, 10.0, 77)
    ao = synth.eri_packed_torch(B, "cuda")
    c = synth.orthonormal(n, 78)[:, :m]
    c_dev = torch.as_tensor(np.ascontiguousarray(c), device="cuda")
    ref = ctx.transform(ao, c_dev).cpu().numpy()
    ao_np = ao.cpu().numpy()
    assert np.array_equal(ctx.transform(ao_np, c), ref)
    # the same output array twice (its pages already mapped), NaN before each call
    buf = np.full(ref.size, np.nan)
    for _ in range(2):
        ctx.transform(ao_np, c, out=buf)
        assert np.array_equal(buf, ref)
        buf[:] = np.nan
    xx, ix = formats.labels_from_packed(ao_np, n)
    assert xx.nbytes > 4 * (64 << 20)
    assert np.array_equal(ctx.itijkl(xx, ix, c), ref)
    ctx.itijkl(xx, ix, c, out=buf)
    assert np.array_equal(buf, ref)
