"""GPU parity: the windowed FFT bandpass against the FP64 oracle with the
reference's windows (chunk 120000, overlap 60000, n = 240000 by the
four-step FFT; other window lengths by the chirp-z form)."""
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
    (72, 150_000, np.int16, 8),        # whole-line int16 staging; the second unit holds 4 of 32 pairs
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


@pytest.mark.parametrize("M,N,dtype,T,cs,k,q", [
    (4, 400_000, np.float32, 8, None, 1, 0.0),      # the 4-channel 8-thread plan: n = 1398101
    (3, 250_000, np.int16, 8, 100_003, 1, 0.0),     # odd channels, several windows, n = 220003
    (5, 260_000, np.float32, 8, 77_777, 3, 0.5),    # subsample 3, int16 output, n = 197777
])
def test_bandpass_any_window_length(gpu, M, N, dtype, T, cs, k, q):
    """Window lengths without a four-step factorization go through the
    chirp-z (Bluestein) form: same windows, within 1e-4 of the FP64 oracle."""
    from mountainsort_amd import frontend as fe
    n = (cs or fe.bandpass_chunk_size(M, T)) + 120000
    say(f"bandpass (chirp-z) M={M} N={N} n={n} k={k} q={q}")
    X = _signal(N, M, seed=M + 7, dtype=dtype)
    Yg = fe.bandpass_filter(torch.from_numpy(X).to(gpu), 30000, 300, 6000, 1000, num_threads=T,
                            chunk_size=cs, subsample_factor=k, quantization_unit=q).cpu().numpy()
    Yo = oc.bandpass_filter(X, 30000, 300, 6000, 1000, num_threads=T, chunk_size=cs,
                            subsample_factor=k, quantization_unit=q)
    assert Yg.shape == Yo.shape
    if q:
        d = np.abs(Yg.astype(np.int32) - Yo.astype(np.int32))
        say(f"  int16 outputs: max diff {d.max()}, fraction differing {np.mean(d > 0):.2e}")
        assert d.max() <= 1 and np.mean(d > 0) < 1e-3
    else:
        err = _rel_err(Yg.astype(np.float64), Yo)
        say(f"  max |y_gpu - y_fp64| / rms = {err:.3e}")
        assert err < 1e-4


def test_bandpass_too_long_window_is_an_error(gpu):
    from mountainsort_amd import frontend as fe
    from mountainsort_amd._capi import MsamdError
    X = torch.zeros((1000, 4), device=gpu)
    with pytest.raises(MsamdError):
        fe.bandpass_filter(X, 30000, 300, 6000, 1000, chunk_size=1 << 25)
