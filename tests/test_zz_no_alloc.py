"""GPU: abm_gini_i32_noalloc (scratch in the library's device array, no
allocation) against abm_gini_i32 (caller's workspace), through the C-ABI:
bit-identical results on two calls in a row, the second on a smaller input,
and a clear error above ABM_GINI_NOALLOC_MAX_N."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_N = 262144


@pytest.fixture(scope="module")
def abi():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from jaxabm_amd import _lib

    return _lib


def _gini_ws(abi, x):
    import torch

    L = abi.lib()
    n = x.numel()
    ws = torch.empty(L.abm_reduce_workspace(n), dtype=torch.uint8, device=x.device)
    out = torch.full((1,), -7.0, dtype=torch.float32, device=x.device)
    abi.check(L.abm_gini_i32(abi.ptr(x), n, abi.ptr(out), abi.ptr(ws), ws.numel(),
                             abi.stream_handle()), "abm_gini_i32")
    return out


def _gini_noalloc(abi, x):
    import torch

    out = torch.full((1,), -7.0, dtype=torch.float32, device=x.device)
    abi.check(abi.lib().abm_gini_i32_noalloc(abi.ptr(x), x.numel(), abi.ptr(out),
                                             abi.stream_handle()), "abm_gini_i32_noalloc")
    return out


def test_noalloc_gini_equals_workspace_gini_on_consecutive_calls(abi):
    import torch

    rs = np.random.default_rng(12)
    # first call at the maximum size, second smaller: the second must not see
    # anything the first left in the shared scratch
    for n, hi in ((MAX_N, 50), (1001, 7), (MAX_N - 3, 2_000_000), (1, 5)):
        x = torch.from_numpy(rs.integers(0, hi, n).astype(np.int32)).cuda()
        a = _gini_noalloc(abi, x)
        b = _gini_ws(abi, x)
        torch.cuda.synchronize()
        assert a.view(torch.int32).item() == b.view(torch.int32).item(), (n, a.item(), b.item())
    # values that wrap the reference's int32 sums, negative values and zeros
    for vals in (np.array([2**30, 2**30, 5, -3, 0, 9], np.int32), np.zeros(17, np.int32)):
        x = torch.from_numpy(vals).cuda()
        a, b = _gini_noalloc(abi, x), _gini_ws(abi, x)
        torch.cuda.synchronize()
        assert a.view(torch.int32).item() == b.view(torch.int32).item()


def test_noalloc_gini_refuses_sizes_above_the_maximum(abi):
    import torch

    x = torch.zeros(MAX_N + 1, dtype=torch.int32, device="cuda")
    out = torch.zeros(1, dtype=torch.float32, device="cuda")
    rc = abi.lib().abm_gini_i32_noalloc(abi.ptr(x), MAX_N + 1, abi.ptr(out), abi.stream_handle())
    assert rc != 0
    assert b"ABM_GINI_NOALLOC_MAX_N" in abi.lib().abm_last_error()
    with pytest.raises(abi.AbmError):
        abi.check(rc, "abm_gini_i32_noalloc")
