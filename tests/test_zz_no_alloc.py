"""pc_nearest_noalloc (the allocation-free variant of pc_nearest, whose
per-query scratch is a static device array of the library) against
pc_nearest with a caller workspace, through the C ABI: identical indices and
distances on two calls in a row, the second with fewer queries (the static
scratch then holds the first call's words and must be cleared by the call),
and sizes beyond the static scratch refused before any GPU work."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _both(L, q, t):
    nq, nt = q.shape[0], t.shape[0]
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    work = torch.empty(nq, dtype=torch.int64, device="cuda")
    i1 = torch.full((nq,), -7, dtype=torch.int32, device="cuda")
    d1 = torch.full((nq,), -7.0, dtype=torch.float32, device="cuda")
    i2 = torch.full((nq,), -9, dtype=torch.int32, device="cuda")
    d2 = torch.full((nq,), -9.0, dtype=torch.float32, device="cuda")
    print("case pc_nearest vs pc_nearest_noalloc: %d queries, %d targets" % (nq, nt), flush=True)
    assert L.pc_nearest(_ptr(q), nq, _ptr(t), nt, _ptr(work), _ptr(i1), _ptr(d1), s) == 0
    assert L.pc_nearest_noalloc(_ptr(q), nq, _ptr(t), nt, _ptr(i2), _ptr(d2), s) == 0, L.pc_last_error()
    torch.cuda.synchronize()
    assert torch.equal(i1, i2)
    assert torch.equal(d1, d2)
    assert int(i2.min()) >= 0 and int(i2.max()) < nt
    return i2, d2


def test_noalloc_nearest_matches_workspace_entry_point():
    from diffskill_amd import _lib
    L = _lib.load()
    g = torch.Generator().manual_seed(0)
    q = torch.rand((6000, 3), generator=g).cuda()
    t = torch.rand((9000, 3), generator=g).cuda()
    i_a, _ = _both(L, q, t)
    # second call, smaller: the scratch holds 6000 words of the first call
    q2 = (torch.rand((1500, 3), generator=g) * 0.5).cuda()
    i_b, d_b = _both(L, q2, t[:2500].contiguous())
    # brute force for the smaller case
    dd = ((q2[:, None, :] - t[None, :2500, :]) ** 2).sum(-1)
    best = dd.min(dim=1).values
    chosen = dd.gather(1, i_b.long()[:, None])[:, 0]
    assert bool((chosen <= best * (1 + 1e-5) + 1e-12).all())
    assert torch.allclose(d_b, best, rtol=1e-5, atol=1e-12)
    # beyond the static scratch: refused on the host, nothing launched
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = L.pc_nearest_noalloc(_ptr(q), (1 << 20) + 1, _ptr(t), 10, _ptr(i_a), _ptr(d_b), s)
    assert rc == -1 and b"scratch" in L.pc_last_error()
