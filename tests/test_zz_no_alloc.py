"""The FP64 transform entry point that allocates nothing
(cipsi_ijkl_transform_packed_noalloc_dev), called through the C ABI and checked against
cipsi_ijkl_transform_packed_dev, which runs the same kernels with the context's own
workspace: the packed MO streams must be identical bit for bit.  Two calls in a row, the
second smaller (the static workspace keeps the first call's contents, which must not
matter), then a size above the static workspace's limit, then the call captured in a
graph and replayed.  The first call of a small size is also checked against the oracle."""
import ctypes

import numpy as np
import pytest

import oracle
from cipsi_amd import synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MAX_AO = 48


@pytest.fixture(scope="module")
def ctx():
    from cipsi_amd import ijkl
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = ijkl.Context(0)
    yield c
    c.close()


def inputs(n, m, seed):
    case = synth.make_case(n, m, 2 * n, 5.0, seed)
    ao = torch.as_tensor(case["ao_packed"], device="cuda")
    c = torch.as_tensor(np.ascontiguousarray(case["c"]), device="cuda")
    return case, ao, c


def test_noalloc_matches_workspace_path_two_calls(ctx):
    for n, m, seed in ((MAX_AO, 44, 901), (20, 13, 902)):     # second call smaller
        print(f"[noalloc] n={n} m={m}", flush=True)
        case, ao, c = inputs(n, m, seed)
        ref = ctx.transform(ao, c)                             # context workspace path
        out = ctx.transform_noalloc(ao, c)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), f"n={n} m={m}"
        if n <= 20:
            o_ref = oracle.ijkl(case["ao_packed"], case["c"])
            assert float(np.max(np.abs(out.cpu().numpy() - o_ref))) <= 1e-12


def test_noalloc_direct_c_abi_and_limits(ctx):
    """Raw ctypes call of the exported symbol; sizes beyond the static workspace are
    refused with CIPSI_ERR_UNSUPPORTED, bad dimensions with CIPSI_ERR_ARG."""
    lib = ctx.lib
    n, m = 31, 30
    print(f"[noalloc-abi] n={n} m={m}", flush=True)
    case, ao, c = inputs(n, m, 903)
    ref = ctx.transform(ao, c)
    ct = c.t().contiguous()
    out = torch.empty_like(ref)
    ctx._use_torch_stream()
    rc = lib.cipsi_ijkl_transform_packed_noalloc_dev(ctx.handle, n, m, ctypes.c_void_p(ao.data_ptr()),
                                                     ctypes.c_void_p(ct.data_ptr()), n,
                                                     ctypes.c_void_p(out.data_ptr()))
    assert rc == 0
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    big = MAX_AO + 1
    dummy = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = ctypes.c_void_p(dummy.data_ptr())
    assert lib.cipsi_ijkl_transform_packed_noalloc_dev(ctx.handle, big, big, p, p, big, p) == -7
    assert lib.cipsi_ijkl_transform_packed_noalloc_dev(ctx.handle, 10, 11, p, p, 10, p) == -1
    assert lib.cipsi_ijkl_transform_packed_noalloc_dev(ctx.handle, 10, 10, p, p, 9, p) == -1
    assert lib.cipsi_ijkl_transform_packed_noalloc_dev(None, 10, 10, p, p, 10, p) == -1


def test_noalloc_under_graph_capture(ctx):
    """Captured with no warm-up call at all (nothing to allocate) and replayed with new
    inputs and eager work in between; every replay equals the workspace path."""
    n, m = 40, 36
    print(f"[noalloc-graph] n={n} m={m}", flush=True)
    cases = [inputs(n, m, 910 + k) for k in range(2)]
    refs = [ctx.transform(ao, c) for _, ao, c in cases]
    _, ao, c = inputs(n, m, 910)
    out = torch.empty_like(refs[0])
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ctx.transform_noalloc(ao, c, out=out)
    for rep in range(2):
        for k, (_, ao_k, c_k) in enumerate(cases):
            ao.copy_(ao_k)
            c.copy_(c_k)
            junk = torch.randn(1 << 18, device="cuda", dtype=torch.float64)
            junk.mul_(3.0)
            out.zero_()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(out, refs[k]), f"replay {rep} case {k}"
