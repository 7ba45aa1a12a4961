"""GPU parity: the windowed FFT bandpass against the FP64 oracle with the
reference's windows (chunk 120000, overlap 60000, n = 240000)."""
import numpy as np
import pytest
import torch

from oracle import core as oc
from tests.conftest import say

pytestmark = pytest.mark.gpu


def _signal(N, M, seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    white = rng.normal(size=(N, M))
    nf = 1 << int(np.ceil(np.log2(N)))
    F = np.fft.rfft(rng.normal(size=(nf, M)), axis=0)
    f = np.arange(F.shape[0], dtype=np.float64)
    f[0] = 1
    pink = np.fft.irfft(F / np.sqrt(f)[:, None], n=nf, axis=0)[:N]
    pink /= pink.std()
    X = (white + pink) * 10 / np.sqrt(2)
    t = np.arange(50)
    for st in rng.integers(0, N - 50, N // 2000):
        X[st:st + 50] += np.outer(-rng.uniform(50, 200) * np.exp(-0.5 * ((t - 20) / 2.5) ** 2),
                                  rng.uniform(0.2, 1, M))
    if dtype == np.int16:
        return np.round(X / 0.195).astype(np.int16)
    return X.astype(np.float32)


def _rel_err(Yg, Yo):
    rms = np.sqrt(np.mean(Yo ** 2, axis=0))
    return np.max(np.abs(Yg - Yo) / rms[None, :])


@pytest.mark.parametrize("M,N,dtype,T", [
    (64, 400_000, np.float32, 8),      # natural plan: chunk 120000
    (7, 250_000, np.int16, 32),        # odd channel count, int16 input, partial last window
    (384, 130_000, np.int16, 1),       # wide probe, two windows
])
def test_bandpass_fp32_within_1e4_of_fp64(gpu, M, N, dtype, T):
    from mountainsort_amd import frontend as fe
    say(f"bandpass M={M} N={N} dtype={np.dtype(dtype).name} T={T}")
    X = _signal(N, M, seed=M, dtype=dtype)
    cs = oc.bandpass_chunk_size(M, T)
    assert cs == fe.bandpass_chunk_size(M, T) == 120000
    Yo = oc.bandpass_filter(X, 30000, 300, 6000, 1000, num_threads=T)
    Yg = fe.bandpass_filter(torch.from_numpy(X).to(gpu), 30000, 300, 6000, 1000,
                            num_threads=T).cpu().numpy()
    assert Yg.shape == Yo.shape and Yg.dtype == np.float32
    err = _rel_err(Yg.astype(np.float64), Yo)
    say(f"  max |y_gpu - y_fp64| / rms = {err:.3e}")
    assert err < 1e-4


def test_bandpass_subsample_and_quantize(gpu):
    from mountainsort_amd import frontend as fe
    say("bandpass subsample 3 + quantization 0.5")
    X = _signal(260_000, 6, seed=5)
    kw = dict(num_threads=1, chunk_size=120000)
    Yo = oc.bandpass_filter(X, 30000, 300, 6000, 1000, subsample_factor=3, **kw)
    Yg = fe.bandpass_filter(torch.from_numpy(X).to(gpu), 30000, 300, 6000, 1000,
                            subsample_factor=3, **kw).cpu().numpy()
    assert Yg.shape == (260_000 // 3, 6)
    assert _rel_err(Yg.astype(np.float64), Yo) < 1e-4
    Qo = oc.bandpass_filter(X, 30000, 300, 6000, 1000, quantization_unit=0.5, **kw)
    Qg = fe.bandpass_filter(torch.from_numpy(X).to(gpu), 30000, 300, 6000, 1000,
                            quantization_unit=0.5, **kw).cpu().numpy()
    assert Qg.dtype == np.int16
    d = np.abs(Qg.astype(np.int32) - Qo.astype(np.int32))
    say(f"  int16 outputs: max diff {d.max()}, fraction differing {np.mean(d > 0):.2e}")
    assert d.max() <= 1 and np.mean(d > 0) < 1e-3


def test_bandpass_unsupported_length_is_an_error(gpu):
    from mountainsort_amd import frontend as fe
    from mountainsort_amd._capi import MsamdError
    X = torch.zeros((1000, 4), device=gpu)
    with pytest.raises(MsamdError):
        fe.bandpass_filter(X, 30000, 300, 6000, 1000, chunk_size=1278101)
