"""Pins the CPU oracle (oracle/) against closed forms, invariants, library
routines and brute force.  The reference ships no tests, goldens or fixtures
for this path (its pca_unit_test only prints), so every check here is chosen
so that a plausible slip in the restatement (a dropped term, a wrong sign or
index, an off-by-one window, a transposed operand) fails.  CPU only."""
import math

import numpy as np
import pytest
from scipy.special import erf

from oracle import core as oc


# ---------------------------------------------------------------- a1 gain
def _gain_ref(n, fs, fmin, fmax, fwid):
    i = np.arange(n)
    df = 1.0 / (n / fs)
    f = np.where(i <= (n + 1) // 2, df * i, df * (i - n))
    a = np.abs(f)
    val = np.ones(n)
    if fmin != 0:
        val = val * (1 + erf(3.0 * (a - fmin) / fmin)) / 2
        val[0] = 0.0
    if fmax != 0:
        val = val * (1 - erf((a - fmax) / fwid)) / 2
    return np.sqrt(val)


@pytest.mark.parametrize("n", [7, 8, 1000, 240000])
def test_define_kernel_closed_form(n):
    H = oc.define_kernel(n, 30000.0, 300.0, 6000.0, 1000.0)
    R = _gain_ref(n, 30000.0, 300.0, 6000.0, 1000.0)
    # scipy's erf and the C library's differ by ulps deep in the erf tails
    np.testing.assert_allclose(H, R, rtol=1e-10, atol=1e-300)
    assert H[0] == 0.0
    # real-to-real symmetry H[i] == H[n-i] except at the fgrid branch point
    assert np.allclose(H[1:], H[1:][::-1], rtol=1e-12) or n % 2 == 1


def test_define_kernel_design_points():
    n = 240000
    fs = 30000.0
    H = oc.define_kernel(n, fs, 300.0, 6000.0, 1000.0)
    df = fs / n
    # -3 dB exactly at freq_min and freq_max (erf(0) = 0), unity in band
    assert abs(H[int(300 / df)] - math.sqrt(0.5) * math.sqrt((1 - math.erf(-5.7)) / 2)) < 1e-12
    assert abs(H[int(6000 / df)] - math.sqrt(0.5) * math.sqrt((1 + math.erf(3 * 19.0)) / 2)) < 1e-12
    # mid-band: the low edge is saturated, the 1000 Hz-wide high edge sits at erf(-3)
    assert abs(H[int(3000 / df)] - math.sqrt((1 - math.erf(-3.0)) / 2)) < 1e-12
    # bin 1 keeps the erf tail (the window-dependent DC leak)
    assert abs(H[1] - math.sqrt((1 + math.erf(3 * (df - 300) / 300)) / 2
                                * (1 - math.erf((df - 6000) / 1000)) / 2)) < 1e-15
    # freq_min = 0 leaves DC at 1 (no high-pass)
    H0 = oc.define_kernel(n, fs, 0.0, 6000.0, 1000.0)
    assert abs(H0[0] - math.sqrt((1 - math.erf(-6.0)) / 2)) < 1e-15


# ------------------------------------------------------ a2/a3 bandpass
def _bandpass_bruteforce(X, fs, fmin, fmax, fwid, cs, ov):
    """Windowed circular convolution with h = IDFT(H), explicit O(n^2) sums."""
    N, M = X.shape
    n = cs + 2 * ov
    H = _gain_ref(n, fs, fmin, fmax, fwid)
    k = np.arange(n)
    h = np.real(np.exp(2j * np.pi * np.outer(k, k) / n) @ H) / n  # h[t] = (1/n) sum_k H[k] e^{2pi i kt/n}
    out = np.zeros((N, M))
    for t in range(0, N, cs):
        win = np.zeros((n, M))
        for s in range(n):
            tt = t - ov + s
            if 0 <= tt < N:
                win[s] = X[tt]
        for r in range(cs):
            if t + r >= N:
                break
            tau = ov + r
            out[t + r] = sum(h[(tau - s) % n] * win[s] for s in range(n))
    return out


def test_bandpass_bruteforce_tiny():
    rng = np.random.default_rng(0)
    X = rng.normal(size=(53, 2)).astype(np.float32)
    Y = oc.bandpass_filter(X, 1000.0, 30.0, 300.0, 50.0, chunk_size=16, overlap=8)
    R = _bandpass_bruteforce(X.astype(np.float64), 1000.0, 30.0, 300.0, 50.0, 16, 8)
    np.testing.assert_allclose(Y, R, atol=1e-12)


def test_bandpass_grid_tone_and_dc():
    """An on-grid tone filling an interior window is scaled by exactly H[k0];
    a constant is removed exactly (H[0] = 0)."""
    fs, cs, ov = 30000.0, 2000, 1000
    n = cs + 2 * ov
    k0 = 123
    N = 3 * cs
    t = np.arange(N)
    tone = np.cos(2 * np.pi * k0 * t / n)
    X = np.stack([tone, np.full(N, 5.0)], axis=1)
    Y = oc.bandpass_filter(X, fs, 300.0, 6000.0, 1000.0, chunk_size=cs, overlap=ov)
    H = oc.define_kernel(n, fs, 300.0, 6000.0, 1000.0)
    core = slice(cs, 2 * cs)  # middle chunk: its window [cs-ov, 2cs+ov) lies inside the data
    # float32 input rounding of the tone is the only error source
    np.testing.assert_allclose(Y[core, 0], H[k0] * tone.astype(np.float32)[core], atol=2e-7)
    np.testing.assert_allclose(Y[core, 1], 0.0, atol=1e-12)


def test_bandpass_linearity_and_subsample_quantize():
    rng = np.random.default_rng(1)
    A = rng.normal(size=(5000, 3)).astype(np.float32)
    B = rng.normal(size=(5000, 3)).astype(np.float32)
    kw = dict(chunk_size=1500, overlap=400)
    ya = oc.bandpass_filter(A, 30000, 300, 6000, 1000, **kw)
    yb = oc.bandpass_filter(B, 30000, 300, 6000, 1000, **kw)
    yab = oc.bandpass_filter((A.astype(np.float64) + B).astype(np.float32), 30000, 300, 6000, 1000, **kw)
    assert np.max(np.abs(yab - (ya + yb))) < 1e-5
    ys = oc.bandpass_filter(A, 30000, 300, 6000, 1000, subsample_factor=3, **kw)
    assert ys.shape == (5000 // 3, 3)
    # chunk 1500 is a multiple of 3, so the subsampled output is every 3rd sample
    np.testing.assert_allclose(ys, ya[: (5000 // 3) * 3: 3], atol=1e-12)
    yq = oc.bandpass_filter(A * 100, 30000, 300, 6000, 1000, quantization_unit=0.5, **kw)
    yf = oc.bandpass_filter(A * 100, 30000, 300, 6000, 1000, **kw)
    assert yq.dtype == np.int16
    np.testing.assert_array_equal(yq, np.trunc((yf.astype(np.float32) * 2.0).astype(np.float32)))


def test_bandpass_freq_max_zero_is_copy():
    X = np.arange(12, dtype=np.float32).reshape(6, 2)
    np.testing.assert_array_equal(oc.bandpass_filter(X, 30000, 300, 0), X)


# ----------------------------------------------------------- chunk plans
def test_bandpass_chunk_plan_values():
    # p_bandpass_filter.cpp:147-155 evaluated by hand
    assert oc.bandpass_chunk_size(64, 8) == 120000
    assert oc.bandpass_chunk_size(64, 2) + 120000 == 349525
    assert oc.bandpass_chunk_size(4, 8) + 120000 == 1398101
    assert oc.bandpass_chunk_size(384, 1) == 120000


def test_detect_chunk_plan_values():
    # p_mountainsort3.cpp:170-172,507-533 evaluated by hand
    assert oc.detect_chunk_size(384, 9_000_000, 8) == 20000
    assert oc.detect_chunk_size(64, 18_000_000, 8) == 488281
    assert oc.detect_chunk_size(64, 18_000_000, 1) == 61035
    assert oc.detect_chunk_size(4, 1_800_000, 8) == 1_800_000
    assert oc.detect_chunk_size(4, 5000, 8) == 20000


# --------------------------------------------------------- a10 detection
def test_mean_stdev_exact():
    m, s = oc.mean_stdev(np.array([1.0, 2.0, 3.0, 4.0]))
    assert m == 2.5 and abs(s - math.sqrt(5.0 / 3.0)) < 1e-15
    assert oc.mean_stdev(np.array([7.0])) == (7.0, 0.0)
    rng = np.random.default_rng(2)
    x = rng.normal(size=10001)
    m, s = oc.mean_stdev(x)
    assert abs(m - math.fsum(x) / len(x)) < 1e-15
    assert abs(s - np.std(x, ddof=1)) < 1e-12


def test_detect_hand_example():
    # mean 0 by construction; thr small
    x = np.zeros(40)
    x[5], x[8], x[20], x[27], x[33] = -5.0, -6.0, -3.0, -3.0, -2.0
    x += -x.mean()
    ev = oc.detect_events_sm(x, 0.5, 4, -1)
    # 8 beats 5 (within 4 samples); 20 and 27 are 7 apart (both kept); 33 is 6 after 27
    assert list(ev) == [8, 20, 27, 33]
    ev2 = oc.detect_events_sm(x, 0.5, 10, -1)
    # with D=10: 8 (6 beats 5); 20 is 12 after 8 -> kept; 27 ties 20 within 10 -> earlier
    # wins; 33 loses to 27 (lag 6) although 27 itself was suppressed
    assert list(ev2) == [8, 20]


@pytest.mark.parametrize("D", [0, 1, 3, 10, 10.5, 25])
@pytest.mark.parametrize("sign", [-1, 0, 1])
def test_detect_state_machine_equals_closed_form(D, sign):
    rng = np.random.default_rng(int(D * 10) + sign + 5)
    n = 20000
    x = rng.normal(size=n)
    for t0 in rng.integers(0, n - 60, 80):
        x[t0:t0 + 30] += rng.normal(size=30) * 4
    # quantized copy: many exact ties and plateaus
    xq = np.round(x * 4) / 4
    for arr in (x, xq):
        a = oc.detect_events_sm(arr, 2.0, D, sign)
        b = oc.detect_events_closed_form(arr, 2.0, D, sign)
        np.testing.assert_array_equal(a, b)


def _ms3alg_detect_bruteforce(X, thr, D, sign, cs, pad, t1, t2):
    N, M = X.shape
    ts = 0 if t1 < 0 else t1
    te = N - 1 if t2 < 0 else t2
    Nr = te - ts + 1
    out = [[] for _ in range(M)]
    for t in range(0, Nr, cs):
        size = min(cs, Nr - t)
        base = ts + t - pad
        win = np.zeros((size + 2 * pad, M))
        for i in range(size + 2 * pad):
            if 0 <= base + i < N:
                win[i] = X[base + i]
        for m in range(M):
            ev = oc.detect_events_closed_form(win[:, m], thr, D, sign)
            for e in ev:
                if pad <= e < pad + size:
                    out[m].append(ts + t + e - pad)
    return out


@pytest.mark.parametrize("t1,t2", [(-1, -1), (3000, 41000)])
def test_ms3alg_detect_chunks(t1, t2):
    rng = np.random.default_rng(7)
    N, M = 47000, 3
    X = rng.normal(size=(N, M)).astype(np.float32)
    X[::97] -= 6
    counts, times = oc.ms3alg_detect(X, 3.0, 10, -1, chunk_size=20000, t1=t1, t2=t2)
    ref = _ms3alg_detect_bruteforce(X.astype(np.float64), 3.0, 10, -1, 20000, 1000, t1, t2)
    off = np.concatenate([[0], np.cumsum(counts)])
    for m in range(M):
        np.testing.assert_array_equal(times[off[m]:off[m + 1]], ref[m])


# --------------------------------------------------------------- a4-a6
def test_whiten_covariance_vs_fp64():
    rng = np.random.default_rng(3)
    X = (rng.normal(size=(250001, 6)) @ rng.normal(size=(6, 6))).astype(np.float32)
    C = oc.whiten_covariance(X)
    R = oc.covariance_fp64(X)
    assert np.max(np.abs(C - R) / np.abs(R).max()) < 1e-7
    np.testing.assert_array_equal(C, C.T)


def test_whitening_matrix_scalar_identity_case():
    """C = c I.  The first component is the normalized start vector; deflation
    leaves a rounding-level residual which the next normalization rescales to
    a unit vector orthogonal to the deflated space, where G acts as c I, so
    every component ends with sigma = c and W = C diag(c^-1/2) C' = I/sqrt(c)
    for whatever orthonormal basis the rounding picks."""
    M, c = 6, 4.0
    W = oc.whitening_matrix(c * np.eye(M))
    np.testing.assert_allclose(W, np.eye(M) / 2.0, atol=1e-12)
    comp, sig = oc.pca_from_XXt(c * np.eye(M), M)
    v = np.sin(np.arange(1, M + 1, dtype=np.float64))
    np.testing.assert_allclose(comp[:, 0], v / np.linalg.norm(v), atol=1e-15)
    np.testing.assert_allclose(sig, c, rtol=1e-12)
    np.testing.assert_allclose(oc.whitening_matrix(np.array([[9.0]])), [[1.0 / 3.0]], rtol=1e-15)


def test_whitening_matrix_converged_case():
    """Eigenvalues 100x apart: 10 power steps converge to machine precision,
    so W equals the exact inverse square root from eigh."""
    rng = np.random.default_rng(4)
    M = 5
    Q, _ = np.linalg.qr(rng.normal(size=(M, M)))
    lam = 100.0 ** np.arange(M)[::-1]
    C = (Q * lam) @ Q.T
    C = (C + C.T) / 2
    W = oc.whitening_matrix(C)
    w, V = np.linalg.eigh(C)
    R = (V / np.sqrt(w)) @ V.T
    assert np.max(np.abs(W - R)) / np.max(np.abs(R)) < 1e-9
    comp, sig = oc.pca_from_XXt(C, M)
    np.testing.assert_allclose(sig, lam, rtol=1e-9)


def _spiked_covariance(M, seed):
    """C = Q diag(lam) Q' with five separated eigenvalues 1e6, 1e5, 1e4, 1e3,
    1e2 (ratio 0.1: ten power steps converge to 1e-10) above an identity
    block.  On the identity block the restarted power method picks whatever
    orthonormal basis the rounding leaves (as for C = c I), and every such
    basis gives the same W.  Returns (C, C^-1/2)."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(M, M)))
    lam = np.ones(M)
    k = min(5, M)
    lam[:k] = 10.0 ** np.arange(6, 6 - k, -1)
    C = (Q * lam) @ Q.T
    return (C + C.T) / 2, (Q / np.sqrt(lam)) @ Q.T


def test_whitening_matrix_spiked_spectrum():
    """Converged components followed by a degenerate block: the result must
    be the exact inverse square root."""
    C, R = _spiked_covariance(24, 6)
    W = oc.whitening_matrix(C)
    assert np.max(np.abs(W - R)) / np.max(np.abs(R)) < 1e-9


def test_whiten_apply_vs_fp64():
    rng = np.random.default_rng(5)
    M, N = 16, 20000
    X = (rng.normal(size=(N, M)) * 10).astype(np.float32)
    W = np.linalg.inv(np.linalg.cholesky(np.cov(X.T.astype(np.float64)))).T
    W = (W + W.T) / 2
    Y = oc.whiten_apply(X, W)
    Y64 = oc.whiten_apply_fp64(X, W)
    k = np.round(Y.astype(np.float64) / 1e-4)
    # outputs are float(k * 1e-4)
    np.testing.assert_array_equal(Y, (k * 1e-4).astype(np.float32))
    kq = np.floor(Y64 / 1e-4 + 0.5)
    assert np.max(np.abs(k - kq)) <= 1
    assert np.mean(k != kq) < 0.05
    Yf, Yq = oc.whiten_apply(X, W, quantization_unit=0.01)
    np.testing.assert_array_equal(Yf, Y)
    np.testing.assert_array_equal(Yq, (Y.astype(np.float64) / 0.01 + 0.5).astype(np.int32).astype(np.int16))


def test_whiten_apply_uses_column_major_walk():
    """out[m1] = sum_m2 x[m2] W(m2, m1): a non-symmetric W tells the two walks apart."""
    X = np.array([[1.0, 0.0], [0.0, 1.0]], np.float32)
    W = np.array([[1.0, 2.0], [3.0, 4.0]])
    Y = oc.whiten_apply(X, W)
    np.testing.assert_allclose(Y, X @ W, atol=1e-4)


# ----------------------------------------------------------------- a7 mask
def test_mask_out_artifacts_rule():
    rng = np.random.default_rng(6)
    N, M, isz = 2000 * 40 + 777, 4, 2000
    X = rng.normal(size=(N, M)).astype(np.float32)
    X[2000 * 10 + 5: 2000 * 10 + 300, 2] += 80  # interval 10, one channel
    X[100:200, 0] += 90                         # interval 0 (left neighbour is out of range)
    Y, norms, use_it = oc.mask_out_artifacts(X, 6.0, isz)
    nint = N // isz
    ref_norms = np.sqrt((X[:nint * isz].astype(np.float64).reshape(nint, isz, M) ** 2).sum(1))
    np.testing.assert_allclose(norms, ref_norms, rtol=1e-13)
    bad = set()
    for m in range(M):
        mu, sd = ref_norms[:, m].mean(), ref_norms[:, m].std(ddof=1)
        bad |= {i for i in range(nint) if ref_norms[i, m] > mu + sd * 6.0}
    assert bad == {0, 10}
    expect = np.ones(nint + 1, np.int32)
    for i in bad:
        for j in (i - 1, i, i + 1):
            if j >= 0:
                expect[j] = 0
    np.testing.assert_array_equal(use_it, expect)
    for i in range(nint):
        seg = slice(i * isz, (i + 1) * isz)
        np.testing.assert_array_equal(Y[seg], X[seg] if expect[i] else 0)
    assert not Y[nint * isz:].any()
    with pytest.raises(ValueError):
        oc.mask_out_artifacts(X, 0.0, isz)


# ------------------------------------------------------------- a11 clips
def test_extract_clips_closed_form():
    X = np.arange(60, dtype=np.float32).reshape(20, 3)
    clips = oc.extract_clips(X, [-2, 5, 18], [2, 0], 4)
    assert clips.shape == (3, 4, 2)
    assert clips[0, 0].tolist() == [0, 0] and clips[0, 2].tolist() == [X[0, 2], X[0, 0]]
    np.testing.assert_array_equal(clips[1], X[5:9][:, [2, 0]])
    assert clips[2, 2].tolist() == [0, 0]
    assert oc.clip_tmid(50) == 24 and oc.clip_tmid(51) == 25
    np.testing.assert_array_equal(oc.extract_clips_processor_starts([100.7, 10, 10.5], 50),
                                  [76, -14, -13])


def test_channels_from_geom():
    geom = np.array([[0, 0], [0, 20], [20, 0], [0, 60]], np.float64)
    assert oc.channels_from_geom(geom, 1, 25) == [1, 2, 3]
    assert oc.channels_from_geom(geom, 4, 25) == [4]
    assert oc.channels_from_geom(geom, 2, -1) == [2, 1, 3, 4]
    assert oc.channels_from_geom(np.zeros((4, 2)), 3, 5) == [3, 1, 2, 4]


# ---------------------------------------------------------------- a12-a15
def _lowrank(rng, M, N, amps, noise=1e-3):
    U, _ = np.linalg.qr(rng.normal(size=(M, len(amps))))
    S = rng.normal(size=(len(amps), N))
    return ((U * amps) @ S + noise * rng.normal(size=(M, N))).T.astype(np.float32), U


def test_pca_matches_svd_when_separated():
    rng = np.random.default_rng(8)
    X, U = _lowrank(rng, 12, 3000, [100.0, 30.0, 9.0, 2.7])
    C, F, s = oc.pca(X, 4)
    Us, sv, _ = np.linalg.svd(X.T.astype(np.float64), full_matrices=False)
    for k in range(4):
        assert abs(abs(np.dot(C[:, k], Us[:, k])) - 1) < 1e-5
        assert abs(s[k] - sv[k] ** 2) / sv[k] ** 2 < 1e-4
    np.testing.assert_allclose(F, X.astype(np.float64) @ C.astype(np.float64), rtol=1e-4, atol=1e-3)


def test_pca_float_form_vs_fp64_gram():
    rng = np.random.default_rng(9)
    X = (rng.normal(size=(4000, 50)) @ rng.normal(size=(50, 50))).astype(np.float32)
    C, F, s = oc.pca_subsampled(X, 10, 10000)
    Cg, Fg, sg = oc.pca_gram_fp64(X, 10, 10000)
    scale = np.abs(Fg).max()
    assert np.max(np.abs(F - Fg)) / scale < 5e-5
    np.testing.assert_allclose(s, sg, rtol=5e-5)


def test_pca_subsample_indices_rule():
    idx = oc.pca_subsample_indices(26106, 10000)
    inc = 26106 / 10000
    acc = np.cumsum(np.full(20000, inc))  # sequential double accumulation i += inc
    ref = np.concatenate([[0], acc[acc < 26106]]).astype(np.int64)
    np.testing.assert_array_equal(idx, ref)
    assert len(oc.pca_subsample_indices(844, 10000)) == 844
    assert len(idx) in (10000, 10001)


def test_dimension_reduce_clips_per_channel():
    rng = np.random.default_rng(10)
    clips = rng.normal(size=(300, 50, 3)).astype(np.float32)
    red = oc.dimension_reduce_clips(clips, 10, 10000)
    assert red.shape == (300, 10, 3)
    for w in range(3):
        _, F, _ = oc.pca_subsampled(clips[:, :, w], 10, 10000)
        np.testing.assert_array_equal(red[:, :, w], F)
    f2 = oc.stage2_features(red, 10, 10000)
    _, F2, _ = oc.pca_subsampled(red.reshape(300, 30), 10, 10000)
    np.testing.assert_array_equal(f2, F2)


def test_two_stage_fp64_form_matches_float_form():
    """The FP64 Gram form of the two-stage PCA (the precision reference of
    the GPU features) agrees with the reference's float data form on clips
    whose spectrum is well separated, and its stage 2 is pca_gram_fp64 of
    the reduced clips viewed as (nf1 * W) x L."""
    rng = np.random.default_rng(12)
    L, T, W = 3000, 50, 4
    t = np.arange(T)
    base = np.stack([np.sin(t / (3.0 + k)) * (8.0 / (k + 1)) for k in range(6)])  # 6 x T
    coef = rng.normal(size=(L, 6, W))
    clips = (np.einsum("lkw,kt->ltw", coef, base) + 0.05 * rng.normal(size=(L, T, W))).astype(np.float32)
    r_o = oc.dimension_reduce_clips(clips, 3, 10000)
    r_64 = oc.dimension_reduce_clips_fp64(clips, 3, 10000)
    assert r_64.shape == r_o.shape
    assert np.max(np.abs(r_o - r_64)) / np.max(np.abs(r_64)) < 1e-4
    f_64 = oc.stage2_features_fp64(r_64, 3, 10000)
    _, F, _ = oc.pca_gram_fp64(r_64.reshape(L, -1), 3, 10000)
    np.testing.assert_array_equal(f_64, F)
    f_o = oc.stage2_features(r_o, 3, 10000)
    assert np.max(np.abs(f_o - f_64)) / np.max(np.abs(f_64)) < 1e-4


# ------------------------------------------- templates and the fit stage
def _spiky(rng, N, M, T, K, n_ev, noise=0.5):
    X = (rng.normal(size=(N, M)) * noise).astype(np.float32)
    t = np.arange(T)
    waves = [(-rng.uniform(3, 6) * np.exp(-0.5 * ((t - 24) / 2.0) ** 2))[:, None]
             * rng.uniform(0.1, 1.0, M)[None, :] for _ in range(K)]
    times = np.sort(rng.choice(np.arange(60, N - 60), n_ev, replace=False))
    labels = rng.integers(1, K + 1, n_ev).astype(np.int32)
    for t0, k in zip(times, labels):
        X[t0 - 24:t0 + 26] += waves[k - 1].astype(np.float32)
    return X, times, labels


def test_compute_templates_is_the_label_mean():
    rng = np.random.default_rng(20)
    N, M, T, K = 5000, 3, 50, 3
    X, times, labels = _spiky(rng, N, M, T, K, 60)
    times = np.concatenate([times, [5, N - 3]])           # clips running off both ends
    labels = np.concatenate([labels, [1, 2]]).astype(np.int32)
    labels[3] = 0                                           # unlabeled: skipped
    tm, cnt = oc.compute_templates(X, times, labels, K, T)
    Xp = np.concatenate([np.zeros((T, M)), X.astype(np.float64), np.zeros((T, M))])
    for k in range(1, K + 1):
        sel = times[labels == k]
        clips = np.stack([Xp[s - 24 + T:s + 26 + T] for s in sel])
        assert cnt[k - 1] == len(sel)
        np.testing.assert_allclose(tm[k - 1], clips.mean(0), rtol=1e-6, atol=1e-6)


def _fit_stage_plain(X, times, labels, templates, thresh=0.1):
    """The fit stage restated without the dirty-flag cache: every considered
    event's score is recomputed on every pass (equivalent, since the cache
    only skips events none of whose masked samples changed)."""
    X = X.astype(np.float32).copy()
    N2, M = X.shape
    K, T = templates.shape[:2]
    masks = [np.nonzero(np.abs(templates[k].astype(np.float64).ravel()) > thresh)[0]
             for k in range(K)]
    consider = list(range(len(times)))
    used = set()
    while True:
        tri = []
        for i in consider:
            k = int(labels[i])
            if k <= 0:
                continue
            tt = int(times[i] - 24 + 0.5)
            s = 0.0
            if tt >= 0 and tt + T <= N2:
                xv = X[tt:tt + T].ravel()
                tv = templates[k - 1].ravel()
                b = a = 0.0
                for j in masks[k - 1]:
                    v = float(xv[j])
                    b += v * v
                    v -= float(tv[j])
                    a += v * v
                s = b - a
            if s > 0:
                tri.append((i, s, float(times[i])))
        if not tri:
            break
        use = []
        for a_, (i, s, t) in enumerate(tri):
            left = [q for q in tri[:a_] if q[2] >= t - T]
            right = [q for q in tri[a_ + 1:] if q[2] <= t + T]
            use.append(all(q[1] < s for q in left) and all(q[1] <= s for q in right))
        changed = False
        nxt = []
        for (i, s, t), u in zip(tri, use):
            if u:
                k = int(labels[i])
                tt = int(t - 24 + 0.5)
                win = X[tt:tt + T].reshape(-1)
                for j in masks[k - 1]:
                    win[j] = np.float32(float(win[j]) - float(templates[k - 1].ravel()[j]))
                X[tt:tt + T] = win.reshape(T, M)
                used.add(i)
                changed = True
            else:
                nxt.append(i)
        consider = nxt
        if not changed:
            break
    keep = np.zeros(len(times), bool)
    keep[list(used)] = True
    return keep


def test_fit_stage_matches_the_uncached_restatement():
    rng = np.random.default_rng(21)
    N, M, T, K = 6000, 4, 50, 3
    X, times, labels = _spiky(rng, N, M, T, K, 90)
    # false events: noise-only times, and two duplicates of a real event
    extra = np.array([200, 211, 3001, times[10], times[10]])
    times = np.concatenate([times, extra])
    labels = np.concatenate([labels, [1, 2, 3, labels[10], labels[10]]]).astype(np.int32)
    o = np.argsort(times, kind="stable")
    times, labels = times[o], labels[o]
    tm, _ = oc.compute_templates(X, times, labels, K, T)
    keep, passes = oc.fit_stage(X.copy(), times.astype(np.float64), labels, tm)
    ref = _fit_stage_plain(X, times, labels, tm)
    assert np.array_equal(keep, ref)
    assert passes >= 2
    # every planted event survives; of the three copies of event 10 one does
    assert keep.sum() >= 90 - 3
    dup = np.nonzero(times == times[np.searchsorted(times, extra[3])])[0]
    assert keep[dup].sum() == 1


def test_time_chunk_infos_closed_form():
    # p_mountainsort3.cpp:507-533
    assert oc.time_chunk_infos(384, 9_000_000, 8, 8) == [
        (t, min(81380, 9_000_000 - t), 1000) for t in range(0, 9_000_000, 81380)]
    assert oc.time_chunk_infos(4, 60_000, 1, 1) == [(0, 60_000, 1000)]
    assert oc.time_chunk_infos(64, 30_000, 8, 8) == [(0, 20_000, 1000), (20_000, 10_000, 1000)]
    # the detection plan is the same rule with one chunk at least and the
    # neighborhood batch count as the number of simultaneous chunks (:170-191)
    for M, N, T in ((384, 9_000_000, 8), (4, 1_800_000, 8), (64, 18_000_000, 16)):
        nb = -(-M // min(T, M))
        assert oc.detect_chunk_size(M, N, T) == oc.time_chunk_infos(M, N, nb, 1)[0][1]


def gap_split(F, min_side=10, frac=0.3):
    """A deterministic stand-in for isosplit5 in tests: split the events at
    the largest gap, relative to the feature's range, over all features when
    it is a large fraction of that range (labels 1, 2), else one cluster.
    (The PCA does not center, so the first feature is mostly the mean.)"""
    F = np.asarray(F, np.float64)
    L0 = len(F)
    lab = np.ones(L0, np.int64)
    if L0 < 2 * min_side + 2:
        return lab
    best = (frac, -1, -1)
    for f in range(F.shape[1]):
        o = np.argsort(F[:, f], kind="stable")
        xs = F[o, f]
        rng_ = xs[-1] - xs[0]
        if not rng_ > 0:
            continue
        gaps = np.diff(xs)[min_side - 1:L0 - min_side] / rng_
        j = int(np.argmax(gaps))
        if gaps[j] > best[0]:
            best = (gaps[j], f, j + min_side - 1)
    if best[1] >= 0:
        o = np.argsort(F[:, best[1]], kind="stable")
        lab[o[best[2] + 1:]] = 2
    return lab


def _clustered_reduced(rng, L_per, K, T=10, M=4, sep=25.0):
    centers = rng.normal(size=(K, T, M)) * sep
    truth = np.repeat(np.arange(K), L_per)
    clips = centers[truth] + rng.normal(size=(K * L_per, T, M))
    o = rng.permutation(len(truth))
    return clips[o].astype(np.float32), truth[o]


def test_sort_clips_branch_method_recovers_separated_clusters():
    rng = np.random.default_rng(30)
    clips, truth = _clustered_reduced(rng, 150, 5)
    lab = oc.sort_clips(clips, gap_split)
    assert lab.min() >= 1
    # every found cluster is pure and every true cluster is found whole
    pairs = {(int(a), int(b)) for a, b in zip(lab, truth)}
    assert len(pairs) == len(set(lab.tolist())) == len(set(truth.tolist())) == 5
