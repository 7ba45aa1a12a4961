"""GPU parity: detection, clips, whitening and masking through the C ABI
against the CPU oracle on the same seeded inputs."""
import numpy as np
import pytest
import torch

from oracle import core as oc
from tests.conftest import say

pytestmark = pytest.mark.gpu


def _recording(M, N, seed, units=6, amp=(60, 200)):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(N, M)) * 10
    t = np.arange(50)
    for _ in range(units):
        w = -rng.uniform(*amp) * np.exp(-0.5 * ((t - 20) / 2.5) ** 2) \
            + 0.3 * rng.uniform(*amp) * np.exp(-0.5 * ((t - 30) / 5) ** 2)
        g = rng.uniform(0.1, 1.0, M)
        for st in rng.integers(0, N - 50, N // 3000):
            X[st:st + 50] += np.outer(w, g)
    return X.astype(np.float32)


def _whitened(M, N, seed):
    """A realistic whitened (1e-4-quantized) recording produced by the oracle."""
    X = _recording(M, N, seed)
    C = oc.whiten_covariance(X)
    W = oc.whitening_matrix(C)
    return oc.whiten_apply(X, W)


# ------------------------------------------------------------------ detect
@pytest.mark.parametrize("M,N,T,sign,D,t1,t2,cs", [
    (4, 200_000, 8, -1, 10, -1, -1, None),
    (4, 200_000, 8, 0, 10, -1, -1, 30_000),
    (7, 150_000, 8, 1, 3, 1234, 120_000, 20_000),
    (64, 100_000, 8, -1, 10, -1, -1, 20_000),
    (64, 100_000, 1, -1, 0, -1, -1, None),
    (3, 45_000, 8, -1, 25, -1, -1, 20_000),
    (7, 150_000, 8, -1, 10, 777, 140_000, 20_000),  # LDS-window stencil at the range's edges
    (6, 60_000, 8, -1, 11, -1, -1, 20_000),   # shortest window past the LDS window (loads)
    (5, 60_000, 8, -1, 16, -1, -1, 20_000),   # longest window of the register/float test
    (5, 60_000, 8, 1, 17, -1, -1, 20_000),    # shortest window of the FP64 neighbour walk
])
def test_detect_bit_exact(gpu, M, N, T, sign, D, t1, t2, cs):
    from mountainsort_amd import frontend as fe
    say(f"detect M={M} N={N} T={T} sign={sign} D={D} t1={t1} t2={t2} cs={cs}")
    Y = _whitened(M, N, seed=M + N)
    n_range = (N - 1 if t2 < 0 else t2) - (0 if t1 < 0 else t1) + 1
    chunk = cs or oc.detect_chunk_size(M, n_range, T)
    assert fe.detect_chunk_size(M, n_range, T) == oc.detect_chunk_size(M, n_range, T)
    counts, times = oc.ms3alg_detect(Y, 3.0, D, sign, chunk_size=chunk, t1=t1, t2=t2)
    off, gt = fe.detect_events(torch.from_numpy(Y).to(gpu), 3.0, D, sign, num_threads=T,
                               chunk_size=chunk, t1=t1, t2=t2)
    np.testing.assert_array_equal(off.cpu().numpy(), np.concatenate([[0], np.cumsum(counts)]))
    np.testing.assert_array_equal(gt.cpu().numpy(), times)
    assert len(times) > 0


@pytest.mark.parametrize("sign", [-1, 1])
def test_detect_rounding_merges(gpu, sign):
    """Excursions to about 1e-3 on channels sitting at 1e8 (opposite to the
    detection sign): x - mean rounds the distinct floats of an excursion to
    the same double, so their order in float is not the order of their
    detection values and only the FP64 test decides them."""
    from mountainsort_amd import frontend as fe
    say(f"detect rounding merges sign={sign}")
    rng = np.random.default_rng(11)
    M, N = 3, 45_000
    Y = (-sign * 1e8 + rng.normal(size=(N, M)) * 100).astype(np.float32)
    for j in range(M):
        for t0 in rng.choice(np.arange(50, N - 50), 40, replace=False):
            run = -sign * (1e-3 + 2.0 ** -30 * rng.integers(0, 4, size=7))
            Y[t0:t0 + 7, j] = run.astype(np.float32)
    counts, times = oc.ms3alg_detect(Y, 3.0, 10, sign, chunk_size=20000)
    off, gt = fe.detect_events(torch.from_numpy(Y).to(gpu), 3.0, 10, sign, chunk_size=20000)
    np.testing.assert_array_equal(off.cpu().numpy(), np.concatenate([[0], np.cumsum(counts)]))
    np.testing.assert_array_equal(gt.cpu().numpy(), times)
    assert len(times) > 0


def test_detect_channel_subset_and_empty(gpu):
    from mountainsort_amd import frontend as fe
    say("detect channel subset / empty range")
    Y = _whitened(6, 60_000, seed=3)
    ch = np.array([4, 1, 5], np.int32)
    counts, times = oc.ms3alg_detect(Y, 3.0, 10, -1, chunk_size=20000, channels=ch)
    off, gt = fe.detect_events(torch.from_numpy(Y).to(gpu), 3.0, 10, -1, chunk_size=20000,
                               channels=ch)
    np.testing.assert_array_equal(gt.cpu().numpy(), times)
    off2, gt2 = fe.detect_events(torch.from_numpy(Y).to(gpu), 3.0, 10, -1, chunk_size=20000,
                                 channels=[])
    assert gt2.numel() == 0 and off2.cpu().tolist() == [0]


# ------------------------------------------------------------------- clips
def test_clips_bit_exact(gpu):
    from mountainsort_amd import frontend as fe
    say("clips single + batched")
    rng = np.random.default_rng(11)
    N, M, T = 30_000, 10, 50
    X = rng.normal(size=(N, M)).astype(np.float32)
    starts = np.concatenate([[-30, -1, 0], rng.integers(0, N - T, 500), [N - 10, N + 5]])
    ch = np.array([3, 0, 9, 4], np.int32)
    Xg = torch.from_numpy(X).to(gpu)
    got = fe.extract_clips(Xg, starts, ch, T).cpu().numpy()
    np.testing.assert_array_equal(got, oc.extract_clips(X, starts, ch, T))
    # batched: three neighborhoods with uneven widths and counts (one empty)
    nb = [np.array([2, 1, 3], np.int32), np.array([5], np.int32), np.array([7, 8, 6, 9, 0], np.int32)]
    ev = [np.sort(rng.integers(0, N, 40)), np.array([], np.int64), np.sort(rng.integers(0, N, 75))]
    times = torch.from_numpy(np.concatenate(ev).astype(np.int64)).to(gpu)
    eo = np.concatenate([[0], np.cumsum([len(e) for e in ev])])
    no = np.concatenate([[0], np.cumsum([len(c) for c in nb])])
    flat, co = fe.extract_clips_batched(Xg, times, eo, np.concatenate(nb), no, T)
    flat = flat.cpu().numpy()
    co = co.cpu().numpy()
    tmid = oc.clip_tmid(T)
    for g in range(3):
        ref = oc.extract_clips(X, ev[g] - tmid, nb[g], T)
        np.testing.assert_array_equal(flat[co[g]:co[g + 1]].reshape(ref.shape), ref)


# --------------------------------------------------------------- whitening
# M covers every kernel instance: the streaming kernel with 5, 10 and 13
# tiles per wave (one or two tile groups) up to M = 384, and the generic
# kernel beyond (385..1024)
@pytest.mark.parametrize("M,N", [(4, 100_000), (16, 50_000), (64, 40_000), (384, 20_000), (5, 1000),
                                 (200, 12_000), (256, 10_000), (300, 9000), (512, 6000),
                                 (1000, 4000)])
def test_covariance_split_bf16_error(gpu, M, N):
    from mountainsort_amd import frontend as fe
    say(f"covariance M={M} N={N}")
    X = _recording(M, N, seed=M)
    C = fe.whiten_covariance(torch.from_numpy(X).to(gpu)).cpu().numpy()
    R = oc.covariance_fp64(X)
    np.testing.assert_array_equal(C, C.T)
    # split-BF16 3-term products, FP32 block sums: relative to the diagonal scale
    err = np.max(np.abs(C - R)) / np.max(np.abs(np.diag(R)))
    say(f"  max |C_gpu - C_fp64| / max diag = {err:.3e}")
    assert err < 2e-5


@pytest.mark.parametrize("M", [1, 3, 4, 8, 16])
def test_whitening_matrix_matches_oracle_small_m(gpu, M):
    from mountainsort_amd import frontend as fe
    say(f"whitening matrix M={M}")
    X = _recording(M, 30_000, seed=100 + M)
    C = oc.whiten_covariance(X)
    Wo = oc.whitening_matrix(C)
    Wg = fe.whitening_matrix(torch.from_numpy(C).to(gpu)).cpu().numpy()
    rel = np.max(np.abs(Wg - Wo)) / np.max(np.abs(Wo))
    say(f"  max rel |W_gpu - W_oracle| = {rel:.3e}")
    assert rel < 1e-10


def test_whitening_matrix_quality_m64(gpu):
    """At M = 64 the 10-step power method is chaotic in rounding; compare the
    whitening quality with the oracle's on the same covariance."""
    from mountainsort_amd import frontend as fe
    say("whitening matrix M=64 quality")
    X = _recording(64, 40_000, seed=164)
    C = oc.whiten_covariance(X)
    Wo = oc.whitening_matrix(C)
    Wg = fe.whitening_matrix(torch.from_numpy(C).to(gpu)).cpu().numpy()
    qo = np.max(np.abs(Wo @ C @ Wo.T - np.eye(64)))
    qg = np.max(np.abs(Wg @ C @ Wg.T - np.eye(64)))
    say(f"  max|WCW'-I| gpu={qg:.3e} oracle={qo:.3e} |dW|/|W|={np.linalg.norm(Wg - Wo) / np.linalg.norm(Wo):.3e}")
    # Rounding-level perturbations of C (1e-14 relative, 30 draws) move the
    # oracle's own max|WCW'-I| over [0.96, 1.09] x its unperturbed value here;
    # the GPU's W must stay within 1.25 x.
    assert qg < 1.25 * qo


@pytest.mark.parametrize("M", [24, 100, 200, 384, 500, 1000])
def test_whitening_matrix_spiked_spectrum(gpu, M):
    """Every kernel instance (16, 32 and 64 lanes per row) at sizes where the
    oracle's chain is chaotic on generic input: on a spiked spectrum W is the
    exact C^-1/2 whatever basis the rounding picks in the identity block."""
    from mountainsort_amd import frontend as fe
    from tests.test_oracle import _spiked_covariance
    say(f"whitening matrix spiked spectrum M={M}")
    C, R = _spiked_covariance(M, 600 + M)
    Wg = fe.whitening_matrix(torch.from_numpy(C).to(gpu)).cpu().numpy()
    rel = np.max(np.abs(Wg - R)) / np.max(np.abs(R))
    say(f"  max rel |W_gpu - C^-1/2| = {rel:.3e}")
    assert rel < 1e-7


@pytest.mark.parametrize("M", [16, 64, 128, 384, 500])
def test_whitening_matrix_xcd_local_exchange_is_bit_identical(gpu, M, monkeypatch):
    """The power method's exchange confined to one XCD (the default where its
    workgroups fit there; M = 500 does not and runs the device-scope form
    either way) gives the same W, bit for bit, as the device-scope exchange
    (MSAMD_W_XCD_LOCAL=0): only where the published words meet changes.  The
    fallback after a failed placement check (forced from the host) gives it
    too."""
    from mountainsort_amd import frontend as fe
    say(f"whitening matrix XCD-local vs device-scope exchange M={M}")
    rng = np.random.default_rng(700 + M)
    A = rng.normal(size=(M, 3 * M))
    C = torch.from_numpy(A @ A.T / (3 * M) + 0.5 * np.eye(M)).to(gpu)
    W_local = [fe.whitening_matrix(C).cpu().numpy() for _ in range(2)]
    monkeypatch.setenv("MSAMD_W_XCD_LOCAL", "0")
    W_dev = fe.whitening_matrix(C).cpu().numpy()
    # the local form told that its workgroups are spread over XCDs stops
    # before any step and the call reruns in the device-scope form
    monkeypatch.setenv("MSAMD_W_XCD_LOCAL", "spread")
    W_fallback = fe.whitening_matrix(C).cpu().numpy()
    assert np.array_equal(W_local[0], W_local[1])
    assert np.array_equal(W_local[0], W_dev)
    assert np.array_equal(W_fallback, W_dev)


@pytest.mark.parametrize("M,N", [(4, 60_000), (16, 30_000), (64, 20_000), (384, 6000), (7, 5000),
                                 (128, 3000), (192, 100_000), (200, 4001), (256, 2500),
                                 (320, 3333), (130, 2000)])
def test_whiten_apply_within_one_quantum(gpu, M, N):
    from mountainsort_amd import frontend as fe
    say(f"whiten apply M={M} N={N}")
    X = _recording(M, N, seed=200 + M)
    C = oc.covariance_fp64(X)
    w, V = np.linalg.eigh(C)
    W = (V / np.sqrt(w)) @ V.T  # any fixed W: the apply is what is tested
    Yg = fe.whiten_apply(torch.from_numpy(X).to(gpu), torch.from_numpy(W).to(gpu)).cpu().numpy()
    k_gpu = np.round(Yg.astype(np.float64) / 1e-4)
    np.testing.assert_array_equal(Yg, (k_gpu * 1e-4).astype(np.float32))
    k64 = np.floor(oc.whiten_apply_fp64(X, W) / 1e-4 + 0.5)
    d = np.abs(k_gpu - k64)
    say(f"  off-by-one-quantum fraction vs FP64: {np.mean(d > 0):.4f}; max {d.max():.0f}")
    assert d.max() <= 1
    Yo = oc.whiten_apply(X, W)
    assert np.max(np.abs(np.round(Yo.astype(np.float64) / 1e-4) - k_gpu)) <= 1
    # the covariance-scaled form (no range pass over X): the same bar
    Cg = torch.from_numpy(C).to(gpu)
    Yc = fe.whiten_apply(torch.from_numpy(X).to(gpu), torch.from_numpy(W).to(gpu),
                         xxt=Cg).cpu().numpy()
    dc = np.abs(np.round(Yc.astype(np.float64) / 1e-4) - k64)
    say(f"  covariance-scaled: off-by-one fraction {np.mean(dc > 0):.4f}; max {dc.max():.0f}")
    assert dc.max() <= 1
    Yq = fe.whiten_apply(torch.from_numpy(X).to(gpu), torch.from_numpy(W).to(gpu),
                         quantization_unit=0.01).cpu().numpy()
    np.testing.assert_array_equal(Yq, (Yg.astype(np.float64) / 0.01 + 0.5).astype(np.int32).astype(np.int16))


# -------------------------------------------------------------------- mask
@pytest.mark.parametrize("M,N,isz", [(4, 2000 * 50 + 123, 2000), (64, 500 * 80, 500), (5, 999, 2000)])
def test_mask_bit_exact(gpu, M, N, isz):
    from mountainsort_amd import frontend as fe
    say(f"mask M={M} N={N} isz={isz}")
    rng = np.random.default_rng(M)
    X = (rng.normal(size=(N, M)) * 10).astype(np.float32)
    for _ in range(3):
        t0 = int(rng.integers(0, max(N - 300, 1)))
        X[t0:t0 + 300, int(rng.integers(0, M))] += 500
    Yo, _, uo = oc.mask_out_artifacts(X, 6.0, isz)
    Yg, ug = fe.mask_out_artifacts(torch.from_numpy(X).to(gpu), 6.0, isz, return_use_it=True)
    np.testing.assert_array_equal(ug.cpu().numpy(), uo)
    np.testing.assert_array_equal(Yg.cpu().numpy(), Yo)


def _spd(rng, M, cond):
    """Random SPD matrix with eigenvalues log-spaced over [1, cond]."""
    Q, _ = np.linalg.qr(rng.normal(size=(M, M)))
    w = np.exp(rng.uniform(0, np.log(cond), M))
    return (Q * w) @ Q.T


@pytest.mark.parametrize("M,cond", [(4, 10.0), (64, 1e3), (130, 1e4), (384, 1e4)])
def test_whitening_matrix_exact_is_inverse_sqrt(gpu, M, cond):
    """Exact mode: W = C^{-1/2} (closed form from numpy's eigh) and W C W = I."""
    from mountainsort_amd import frontend as fe
    say(f"exact whitening matrix M={M} cond={cond:g}")
    rng = np.random.default_rng(M)
    C = _spd(rng, M, cond) * 37.0
    w, V = np.linalg.eigh(C)
    W_ref = (V / np.sqrt(w)) @ V.T
    Wg, (iters, resid) = fe.whitening_matrix(torch.from_numpy(C).to(gpu), exact=True,
                                             return_info=True)
    Wg = Wg.cpu().numpy()
    err = np.linalg.norm(Wg - W_ref) / np.linalg.norm(W_ref)
    say(f"  {iters} iterations, residual {resid:.2e}, |W - C^-1/2| / |C^-1/2| = {err:.2e}")
    assert err < 1e-10
    assert np.allclose(Wg, Wg.T, rtol=0, atol=1e-14 * np.abs(Wg).max())
    assert np.max(np.abs(Wg @ C @ Wg - np.eye(M))) < 1e-9


def test_whitening_matrix_exact_zero_channels_and_singular(gpu):
    """Channels with a zero diagonal get zero rows/columns; a singular
    covariance on the nonzero channels is an error, not a wrong W."""
    from mountainsort_amd import _capi
    from mountainsort_amd import frontend as fe
    rng = np.random.default_rng(9)
    M = 12
    C = np.zeros((M, M))
    act = np.array([0, 1, 2, 4, 5, 7, 8, 9, 11])
    C[np.ix_(act, act)] = _spd(rng, len(act), 100.0)
    Wg = fe.whitening_matrix(torch.from_numpy(C).to(gpu), exact=True).cpu().numpy()
    w, V = np.linalg.eigh(C[np.ix_(act, act)])
    Wa = (V / np.sqrt(w)) @ V.T
    assert np.linalg.norm(Wg[np.ix_(act, act)] - Wa) / np.linalg.norm(Wa) < 1e-10
    inactive = np.setdiff1d(np.arange(M), act)
    assert np.all(Wg[inactive, :] == 0) and np.all(Wg[:, inactive] == 0)
    x = rng.normal(size=(M, 3))
    S = x @ x.T  # rank 3, every diagonal entry nonzero
    with pytest.raises(_capi.MsamdError):
        fe.whitening_matrix(torch.from_numpy(S).to(gpu), exact=True)


@pytest.mark.parametrize("M,N,isz", [(64, 2000 * 37 + 555, 2000), (384, 500 * 41, 500),
                                     (7, 999, 2000), (256, 500 * 23 + 17, 500)])
def test_masked_whitening_without_the_masked_copy(gpu, M, N, isz):
    """mask_rows + the row-bitmap covariance/apply equal mask_out_artifacts
    followed by the plain covariance/apply, bit for bit."""
    from mountainsort_amd import frontend as fe
    say(f"masked whitening by row bitmap M={M} N={N} isz={isz}")
    X = _recording(M, N, seed=300 + M)
    rng = np.random.default_rng(M)
    nint = N // isz
    for i in rng.choice(max(nint, 1), size=min(3, max(nint, 1)), replace=False):
        X[i * isz:(i + 1) * isz] *= 80.0   # artifacts
    Xg = torch.from_numpy(X).to(gpu)
    Xm, use_it = fe.mask_out_artifacts(Xg, 6.0, isz, return_use_it=True)
    rk, use_it2 = fe.mask_rows(Xg, 6.0, isz)
    assert torch.equal(use_it, use_it2)
    bits = rk.cpu().numpy().view(np.uint32)
    t = np.arange(N)
    kept = (bits[t >> 5] >> (t & 31)) & 1
    u = use_it.cpu().numpy()
    expect = (t // isz < nint) & (u[np.minimum(t // isz, nint)] != 0)
    assert np.array_equal(kept.astype(bool), expect)
    C1 = fe.whiten_covariance(Xm)
    C2 = fe.whiten_covariance(Xg, row_keep=rk)
    assert torch.equal(C1, C2)
    w, V = np.linalg.eigh(C1.cpu().numpy() + 1e-3 * np.eye(M))
    W = torch.from_numpy((V / np.sqrt(w)) @ V.T).to(gpu)
    Y1 = fe.whiten_apply(Xm, W)
    Y2 = fe.whiten_apply(Xg, W, row_keep=rk)
    assert torch.equal(Y1, Y2)
    # with the covariance bound instead of the range pass: within one quantum
    Y3 = fe.whiten_apply(Xg, W, row_keep=rk, xxt=C2)
    assert torch.max(torch.abs(torch.round(Y3.double() / 1e-4) - torch.round(Y1.double() / 1e-4))) <= 1


@pytest.mark.parametrize("L,T,M", [(1000, 50, 12), (37, 50, 64), (500, 20, 384), (0, 50, 8)])
def test_whiten_clips_within_fp32_of_fp64(gpu, L, T, M):
    """whiten_clips (no 1e-4 quantize) against the FP64 product and the
    reference's float loop."""
    from mountainsort_amd import frontend as fe
    say(f"whiten_clips L={L} T={T} M={M}")
    rng = np.random.default_rng(L + M)
    clips = (rng.normal(size=(L, T, M)) * 30).astype(np.float32)
    A = rng.normal(size=(M, M)) / np.sqrt(M)
    W = A @ A.T + 0.1 * np.eye(M)
    Y = fe.whiten_clips(torch.from_numpy(clips).to(gpu), torch.from_numpy(W).to(gpu))
    Y = Y.cpu().numpy()
    assert Y.shape == (L, T, M)
    if L == 0:
        return
    Y64 = np.einsum("ltm,km->ltk", clips.astype(np.float64), W)
    rms = np.sqrt(np.mean(Y64 ** 2))
    e = np.max(np.abs(Y - Y64)) / rms
    say(f"  max |y - y_fp64| / rms = {e:.2e}")
    assert e < 1e-5
    ref = oc.whiten_clips(clips, W)
    assert np.max(np.abs(Y - ref)) / rms < 1e-5


@pytest.mark.parametrize("scale", [1e-5, 1e5, 3e7])
def test_whiten_apply_scaling_extremes(gpu, scale):
    """The fp16 split scales both operands by powers of two: data far from
    unit size (W compensating) still whitens within one quantum of FP64."""
    from mountainsort_amd import frontend as fe
    M, N = 96, 20_000
    say(f"whiten apply scale={scale:g}")
    X = (_recording(M, N, seed=77) * scale).astype(np.float32)
    C = oc.covariance_fp64(X)
    w, V = np.linalg.eigh(C)
    W = (V / np.sqrt(w)) @ V.T
    k64 = np.floor(oc.whiten_apply_fp64(X, W) / 1e-4 + 0.5)
    for xxt in (None, torch.from_numpy(C).to(gpu)):
        Y = fe.whiten_apply(torch.from_numpy(X).to(gpu), torch.from_numpy(W).to(gpu),
                            xxt=xxt).cpu().numpy()
        d = np.abs(np.round(Y.astype(np.float64) / 1e-4) - k64)
        say(f"  xxt={'yes' if xxt is not None else 'no'}: off-by-one {np.mean(d > 0):.4f}, max {d.max():.0f}")
        assert d.max() <= 1


def test_whitening_matrix_wait_timeout_is_reported(gpu, monkeypatch):
    """The power method's bounded waits, with their spin limit forced to
    zero from the host: the first wait gives up, every wavefront still runs
    to its end, and the call returns ERR_TIMEOUT with its message;
    fallback='exact' returns the exact form with a warning; the next call
    with the normal limit succeeds and matches a call made before."""
    import torch

    from mountainsort_amd import _capi
    from mountainsort_amd import frontend as fe
    rng = np.random.default_rng(6)
    M = 384
    A = rng.normal(size=(M, 2 * M))
    C = torch.from_numpy(A @ A.T / (2 * M) + np.eye(M)).to(gpu)
    W_before = fe.whitening_matrix(C).cpu().numpy()
    say("whitening matrix, spin limit 0")
    monkeypatch.setenv("MSAMD_W_SPIN_LIMIT", "0")
    with pytest.raises(_capi.MsamdError) as ei:
        fe.whitening_matrix(C)
    assert ei.value.code == _capi.ERR_TIMEOUT and "timed out" in str(ei.value)
    with pytest.warns(RuntimeWarning, match="exact"):
        Wx = fe.whitening_matrix(C, fallback="exact").cpu().numpy()
    assert np.array_equal(Wx, fe.whitening_matrix(C, exact=True).cpu().numpy())
    monkeypatch.delenv("MSAMD_W_SPIN_LIMIT")
    W_after = fe.whitening_matrix(C).cpu().numpy()
    assert np.array_equal(W_after, W_before)
