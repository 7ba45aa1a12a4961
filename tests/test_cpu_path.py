"""The threaded CPU path timed as the CPU baseline (oracle.cpu_path) against
the sequential oracle: the same results from the reference's thread
structure."""
import numpy as np
import pytest

from oracle import core as oc
from oracle import cpu_path as cp


@pytest.fixture(scope="module")
def rec():
    from mountainsort_amd import synth
    cfg = synth.SynthConfig(M=6, N=150_000, num_units=4, seed=11)
    return synth.generate(cfg, device="cpu").numpy().astype(np.float32)


def test_bandpass_windows_match_the_fp64_oracle(rec):
    X = rec[:, :4]
    ref = oc.bandpass_filter(X, 30000, 300, 6000, num_threads=3, chunk_size=40000)
    got = cp.bandpass_mt(X, 30000, 300, 6000, num_threads=3, chunk_size=40000)
    rms = np.sqrt(np.mean(ref ** 2))
    assert got.dtype == np.float32
    assert np.max(np.abs(got - ref)) < 1e-5 * rms
    # a bounded sample: only the listed windows' cores are written
    part = cp.bandpass_mt(X, 30000, 300, 6000, num_threads=2, chunk_size=40000,
                          windows=[40000])
    assert np.array_equal(part[40000:80000], got[40000:80000])
    assert not part[:40000].any() and not part[80000:].any()


def test_whitening_threads_match_the_sequential_oracle(rec):
    X = rec[:5503]
    C0 = oc.whiten_covariance(X)
    C1 = cp.whiten_covariance_mt(X, num_threads=4, chunk_size=1000)
    # the same float products and per-chunk sums, merged in completion order
    assert np.max(np.abs(C1 - C0)) <= 1e-12 * np.max(np.abs(C0))
    W = oc.whitening_matrix(C0)
    Y0 = oc.whiten_apply(X, W)
    Y1 = cp.whiten_apply_mt(X, W, num_threads=4, chunk_size=1000)
    assert np.array_equal(Y1, Y0)


def test_read_phase_matches_detect_and_gather(rec):
    X = oc.quantize_1e4(rec / np.std(rec, axis=0))
    N, M = X.shape
    nb = [[m] + [c for c in range(M) if c != m and abs(c - m) <= 2] for m in range(M)]
    cs = 20000
    clips = cp.ms3alg_clips_mt(X, nb, 3.0, 10, -1, 50, num_threads=4, chunk_size=cs)
    counts, times = oc.ms3alg_detect(X, 3.0, 10, -1, chunk_size=cs)
    off = np.concatenate([[0], np.cumsum(counts)])
    for m in range(M):
        ev = times[off[m]:off[m + 1]]
        ref = oc.extract_clips(X, ev - oc.clip_tmid(50), nb[m], 50)
        assert clips[m].shape == ref.shape
        assert np.array_equal(clips[m], ref)
    assert sum(len(c) for c in clips) > 100


def test_pca_threads_match_the_sequential_oracle():
    rng = np.random.default_rng(4)
    clips = [rng.normal(size=(L, 50, W)).astype(np.float32)
             for L, W in [(300, 3), (0, 2), (12000, 2), (50, 5), (800, 1)]]
    red, feat = cp.pca_mt(clips, num_threads=3, return_results=True)
    for c, r, f in zip(clips, red, feat):
        if len(c) == 0:
            continue
        r0 = oc.dimension_reduce_clips(c, 10, 10000)
        assert np.array_equal(r, r0)
        assert np.array_equal(f, oc.stage2_features(r0, 10, 10000))
