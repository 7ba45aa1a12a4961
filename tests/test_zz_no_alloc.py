"""msamd_whiten_apply_cov_noalloc (scratch in the library's own device array)
against msamd_whiten_apply_cov (scratch in a caller workspace), called
through the C ABI: bit-identical float and int16 outputs on two calls in a
row, the second with fewer channels and samples, and a clean capacity error
beyond 512 channels (refused before any launch)."""
import numpy as np
import pytest
import torch

from mountainsort_amd import _capi
from tests.conftest import say

pytestmark = pytest.mark.gpu


def _case(gpu, M, N, seed, masked):
    rng = np.random.default_rng(seed)
    X = torch.from_numpy((rng.normal(size=(N, M)) * 20).astype(np.float32)).to(gpu)
    A = rng.normal(size=(M, M)) / np.sqrt(M)
    Wcm = torch.from_numpy(np.ascontiguousarray((A @ A.T + 0.5 * np.eye(M)).T)).to(gpu)
    Xd = X.double()
    C = (Xd.t() @ Xd / max(N - 1, 1)).contiguous()
    rk = None
    if masked:
        words = (N + 31) // 32
        bits = rng.integers(0, 2 ** 32, size=words, dtype=np.uint64).astype(np.uint32)
        rk = torch.from_numpy(bits.view(np.int32)).to(gpu)
    return X, Wcm, C, rk


def _p(t):
    return t.data_ptr() if t is not None else None


def test_noalloc_apply_matches_workspace_form(gpu):
    L = _capi.lib()
    stream = torch.cuda.current_stream(gpu).cuda_stream
    for M, N, masked, q in ((384, 50_000, True, 0.0), (64, 20_011, False, 0.0),
                            (384, 30_000, False, 0.05), (7, 5000, True, 0.05)):
        say(f"noalloc apply M={M} N={N} masked={masked} q={q}")
        X, Wcm, C, rk = _case(gpu, M, N, M + N, masked)
        nws = L.msamd_whiten_apply_workspace_size(M)
        ws = torch.empty(nws, dtype=torch.uint8, device=gpu)
        if q > 0:
            ya = torch.empty((N, M), dtype=torch.int16, device=gpu)
            yb = torch.full((N, M), -7, dtype=torch.int16, device=gpu)
            args_a = (None, q, _p(ya))
            args_b = (None, q, _p(yb))
        else:
            ya = torch.empty((N, M), dtype=torch.float32, device=gpu)
            yb = torch.full((N, M), float("nan"), dtype=torch.float32, device=gpu)
            args_a = (_p(ya), 0.0, None)
            args_b = (_p(yb), 0.0, None)
        _capi.call("msamd_whiten_apply_cov", _p(X), M, N, _p(Wcm), _p(C), _p(rk), *args_a,
                   _p(ws), nws, stream)
        _capi.call("msamd_whiten_apply_cov_noalloc", _p(X), M, N, _p(Wcm), _p(C), _p(rk),
                   *args_b, stream)
        torch.cuda.synchronize()
        assert torch.equal(ya, yb)
        assert bool(torch.any(ya != 0))


def test_noalloc_apply_refuses_more_than_512_channels(gpu):
    L = _capi.lib()
    stream = torch.cuda.current_stream(gpu).cuda_stream
    M, N = 513, 64
    X = torch.zeros((N, M), dtype=torch.float32, device=gpu)
    W = torch.eye(M, dtype=torch.float64, device=gpu)
    Y = torch.empty_like(X)
    rc = L.msamd_whiten_apply_cov_noalloc(_p(X), M, N, _p(W), _p(W), None, _p(Y), 0.0, None,
                                          stream)
    assert rc == 4  # MSAMD_ERR_CAPACITY
    assert b"1..512" in L.msamd_last_error()
