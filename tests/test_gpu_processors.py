"""GPU parity of the msamd.mp processors (the MountainLab drop-in boundary):
.mda files in, .mda files out, compared with the CPU oracle on the same
inputs."""
import os
import subprocess

import numpy as np
import pytest

from mountainsort_amd import mdaio
from oracle import core as oc
from tests.conftest import PCA_STAGE2_BOUND, say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mountainsort_amd", "bin", "msamd.mp")


def run(*args, env=None):
    say("msamd.mp", args[0], env or "")
    full = dict(os.environ, **env) if env else None
    r = subprocess.run([EXE, *map(str, args)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=300, env=full)
    assert r.returncode == 0, r.stderr[-2000:]
    return r


def _recording(M, N, seed, units=5, amp=(60, 200)):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(N, M)) * 10
    t = np.arange(50)
    for _ in range(units):
        w = -rng.uniform(*amp) * np.exp(-0.5 * ((t - 20) / 2.5) ** 2) \
            + 0.3 * rng.uniform(*amp) * np.exp(-0.5 * ((t - 30) / 5) ** 2)
        g = rng.uniform(0.1, 1.0, M)
        for st in rng.integers(0, N - 50, N // 3000):
            X[st:st + 50] += np.outer(w, g)
    return X


def _quanta(a, b):
    """Largest difference in units of the 1e-4 output quantum."""
    qa = np.rint(np.asarray(a, np.float64) / 1e-4)
    qb = np.rint(np.asarray(b, np.float64) / 1e-4)
    return int(np.max(np.abs(qa - qb)))


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        pytest.skip("msamd.mp not built")
    return EXE


def test_bandpass_filter_processor(gpu, exe, tmp_path):
    X = np.clip(_recording(8, 300_000, 1), -32768, 32767).astype(np.int16)
    mdaio.writemda(tmp_path / "raw.mda", X)
    run("ms3.bandpass_filter", f"--timeseries={tmp_path / 'raw.mda'}",
        f"--timeseries_out={tmp_path / 'filt.mda'}", "--samplerate=30000", "--freq_min=300",
        "--freq_max=6000", "--_request_num_threads=32")
    Y = mdaio.readmda(tmp_path / "filt.mda")
    ref = oc.bandpass_filter(X, 30000, 300, 6000, num_threads=32)
    assert Y.dtype == np.float32 and Y.shape == ref.shape
    err = np.max(np.abs(Y - ref)) / np.sqrt(np.mean(ref ** 2))
    say(f"  bandpass max err / rms = {err:.2e}")
    assert err < 1e-4


@pytest.mark.parametrize("M,N,threads,k,q,dtype", [
    (8, 1_000_000, 32, 1, 0.0, np.int16),       # four-step windows, int16 in, f32 out
    (6, 700_000, 32, 7, 0.0, np.float32),       # subsample 7 does not divide the chunk
    (5, 600_000, 32, 1, 0.05, np.int16),        # int16 output
    (4, 3_000_000, 8, 1, 0.0, np.float32),      # Bluestein windows (chunk 1278101)
])
def test_bandpass_streamed_segments(gpu, exe, tmp_path, M, N, threads, k, q, dtype):
    """Streaming a few windows per segment from a recording split over two
    files gives the single-segment output bit for bit, and the oracle's
    within the filter bound."""
    X = np.clip(_recording(M, N, 11 + M), -32768, 32767).astype(dtype)
    d = tmp_path / "parts"
    d.mkdir()
    cut = N // 3 + 17
    mdaio.writemda(d / "a.mda", X[:cut])
    mdaio.writemda(d / "b.mda", X[cut:])
    args = [f"--timeseries={d}", "--samplerate=30000", "--freq_min=300", "--freq_max=6000",
            f"--_request_num_threads={threads}", f"--subsample_factor={k}",
            f"--quantization_unit={q}"]
    cs = oc.bandpass_chunk_size(M, threads)
    row = M * X.dtype.itemsize
    run("ms3.bandpass_filter", *args, f"--timeseries_out={tmp_path / 'one.mda'}")
    run("ms3.bandpass_filter", *args, f"--timeseries_out={tmp_path / 'seg.mda'}",
        env={"MSAMD_SEGMENT_BYTES": str(2 * cs * row)})
    one = mdaio.readmda(tmp_path / "one.mda")
    seg = mdaio.readmda(tmp_path / "seg.mda")
    nseg = -(-(-(-N // cs)) // 2)
    say(f"  {M} ch x {N}, chunk {cs}, {nseg} segments of 2 windows")
    assert one.shape == (N // k, M) and np.array_equal(one, seg)
    ref = oc.bandpass_filter(X, 30000, 300, 6000, num_threads=threads, subsample_factor=k,
                             quantization_unit=q)
    if q:
        assert one.dtype == np.int16
        assert np.max(np.abs(one.astype(np.int64) - ref.astype(np.int64))) <= 1
    else:
        err = np.max(np.abs(one - ref)) / np.sqrt(np.mean(ref ** 2))
        say(f"  max err / rms = {err:.2e}")
        assert err < 1e-4


def test_whiten_streamed_segments(gpu, exe, tmp_path):
    """ms3.whiten over row segments: the device-resident and the two-pass
    (re-read) forms agree bit for bit, and both agree with the oracle to one
    quantum; ms3.apply_whitening_matrix streams the same way."""
    M, N = 12, 400_000
    X = _recording(M, N, 21).astype(np.float32)
    mdaio.writemda(tmp_path / "x.mda", X)
    seg = {"MSAMD_SEGMENT_BYTES": str(37_001 * M * 4)}
    run("ms3.whiten", f"--timeseries={tmp_path / 'x.mda'}", f"--timeseries_out={tmp_path / 'r.mda'}",
        env=seg)
    run("ms3.whiten", f"--timeseries={tmp_path / 'x.mda'}", f"--timeseries_out={tmp_path / 't.mda'}",
        env=dict(seg, MSAMD_STREAM_RESIDENT="0"))
    Yr = mdaio.readmda(tmp_path / "r.mda")
    Yt = mdaio.readmda(tmp_path / "t.mda")
    assert np.array_equal(Yr, Yt)
    ref = oc.whiten_apply(X, oc.whitening_matrix(oc.whiten_covariance(X)))
    d = _quanta(Yr, ref)
    say(f"  whiten streamed: max |diff| = {d} quanta of 1e-4")
    assert d <= 1
    rng = np.random.default_rng(5)
    W = rng.normal(size=(M, M)) * 0.1
    mdaio.writemda(tmp_path / "Wn.mda", np.ascontiguousarray(W.T))
    run("ms3.apply_whitening_matrix", f"--timeseries={tmp_path / 'x.mda'}",
        f"--whitening_matrix={tmp_path / 'Wn.mda'}", f"--timeseries_out={tmp_path / 'a.mda'}",
        "--quantization_unit=0.01", env=seg)
    A = mdaio.readmda(tmp_path / "a.mda")
    _, refA = oc.whiten_apply(X, W, quantization_unit=0.01)
    assert A.dtype == np.int16
    assert np.max(np.abs(A.astype(np.int64) - refA.astype(np.int64))) <= 1


def test_mask_processor(gpu, exe, tmp_path):
    X = _recording(6, 100_000, 2)
    X[40_000:41_500] += 3000.0
    X = X.astype(np.float32)
    mdaio.writemda(tmp_path / "x.mda", X)
    run("ms3.mask_out_artifacts", f"--timeseries={tmp_path / 'x.mda'}",
        f"--timeseries_out={tmp_path / 'y.mda'}", "--threshold=6", "--interval_size=2000")
    Y = mdaio.readmda(tmp_path / "y.mda")
    ref = oc.mask_out_artifacts(X, 6, 2000)[0]
    assert np.array_equal(Y, ref)
    assert np.all(Y[40_000:41_500] == 0)


def test_whiten_processors(gpu, exe, tmp_path):
    M, N = 8, 200_000
    X = _recording(M, N, 3).astype(np.float32)
    mdaio.writemda(tmp_path / "x.mda", X)
    # ms3.whiten
    run("ms3.whiten", f"--timeseries={tmp_path / 'x.mda'}", f"--timeseries_out={tmp_path / 'w.mda'}")
    Y = mdaio.readmda(tmp_path / "w.mda")
    Wref = oc.whitening_matrix(oc.whiten_covariance(X))
    ref = oc.whiten_apply(X, Wref)
    d = _quanta(Y, ref)
    say(f"  whiten max |diff| = {d} quanta of 1e-4")
    assert d <= 1
    # ms3.compute_whitening_matrix on a channel subset
    run("ms3.compute_whitening_matrix", f"--timeseries_list={tmp_path / 'x.mda'}",
        f"--whitening_matrix_out={tmp_path / 'W.mda'}", "--channels=3,1,5")
    Wf = mdaio.readmda(tmp_path / "W.mda").T   # [row, col]
    Wsub = oc.whitening_matrix(oc.whiten_covariance(X[:, [2, 0, 4]]))
    e = np.max(np.abs(Wf - Wsub)) / np.max(np.abs(Wsub))
    say(f"  compute_whitening_matrix rel err = {e:.2e}")
    assert Wf.shape == (3, 3) and e < 1e-4
    # ms3.apply_whitening_matrix with a non-symmetric matrix, int16 output
    rng = np.random.default_rng(4)
    W = rng.normal(size=(M, M)) * 0.1
    mdaio.writemda(tmp_path / "Wn.mda", np.ascontiguousarray(W.T))   # column-major W
    run("ms3.apply_whitening_matrix", f"--timeseries={tmp_path / 'x.mda'}",
        f"--whitening_matrix={tmp_path / 'Wn.mda'}", f"--timeseries_out={tmp_path / 'a.mda'}")
    A = mdaio.readmda(tmp_path / "a.mda")
    refA = oc.whiten_apply(X, W)
    d = _quanta(A, refA)
    say(f"  apply_whitening_matrix max |diff| = {d} quanta of 1e-4")
    assert d <= 1


def test_whiten_processor_falls_back_to_exact(gpu, exe, tmp_path):
    """When the power method's waits give up (spin limit forced to zero),
    ms3.whiten retries, then warns and whitens with the exact C^-1/2.  (At
    384 channels the power method runs 96 workgroups, so a wait is certain.)"""
    M, N = 384, 40_000
    X = _recording(M, N, 31).astype(np.float32)
    mdaio.writemda(tmp_path / "x.mda", X)
    r = run("ms3.whiten", f"--timeseries={tmp_path / 'x.mda'}",
            f"--timeseries_out={tmp_path / 'w.mda'}", env={"MSAMD_W_SPIN_LIMIT": "0"})
    assert "exact whitening matrix" in r.stderr
    Y = mdaio.readmda(tmp_path / "w.mda")
    w, V = np.linalg.eigh(oc.covariance_fp64(X))
    Wx = (V / np.sqrt(w)) @ V.T
    d = _quanta(Y, oc.whiten_apply(X, Wx))
    say(f"  exact fallback: max |diff| = {d} quanta of 1e-4")
    assert d <= 1


def test_whiten_clips_processor(gpu, exe, tmp_path):
    """ms3.whiten_clips: W on every time sample of every clip, no 1e-4 quantize;
    float output within FP32 rounding of the reference's float loop, int16
    output within one count."""
    rng = np.random.default_rng(12)
    L, T, M = 300, 50, 9
    clips = (rng.normal(size=(L, T, M)) * 20).astype(np.float32)
    W = rng.normal(size=(M, M)) * 0.1
    W = W + W.T
    mdaio.writemda(tmp_path / "c.mda", clips)              # M x T x L
    mdaio.writemda(tmp_path / "W.mda", np.ascontiguousarray(W.T))
    say(f"whiten_clips processor L={L} T={T} M={M}")
    run("ms3.whiten_clips", f"--clips={tmp_path / 'c.mda'}",
        f"--whitening_matrix={tmp_path / 'W.mda'}", f"--clips_out={tmp_path / 'o.mda'}")
    out = mdaio.readmda(tmp_path / "o.mda")
    ref = oc.whiten_clips(clips, W)
    assert out.shape == ref.shape and out.dtype == np.float32
    e = np.max(np.abs(out - ref)) / np.max(np.abs(ref))
    say(f"  float output: max rel diff to the reference loop {e:.2e}")
    assert e < 1e-5
    run("ms3.whiten_clips", f"--clips={tmp_path / 'c.mda'}",
        f"--whitening_matrix={tmp_path / 'W.mda'}", f"--clips_out={tmp_path / 'q.mda'}",
        "--quantization_unit=0.1")
    q = mdaio.readmda(tmp_path / "q.mda")
    _, refq = oc.whiten_clips(clips, W, quantization_unit=0.1)
    assert q.dtype == np.int16
    assert np.max(np.abs(q.astype(np.int32) - refq.astype(np.int32))) <= 1


def test_extract_clips_processor(gpu, exe, tmp_path):
    X = _recording(5, 50_000, 5).astype(np.float32)
    mdaio.writemda(tmp_path / "x.mda", X)
    et = np.array([0.0, 3.7, 24.0, 1000.2, 25_000.0, 49_990.0, 49_999.0, -3.0])
    mdaio.writemda(tmp_path / "et.mda", et.reshape(1, -1))
    run("ms3.extract_clips", f"--timeseries={tmp_path / 'x.mda'}",
        f"--event_times={tmp_path / 'et.mda'}", f"--clips_out={tmp_path / 'c.mda'}",
        "--clip_size=50", "--channels=2,4")
    C = mdaio.readmda(tmp_path / "c.mda")   # (L, T, W)
    ref = oc.extract_clips(X, oc.extract_clips_processor_starts(et, 50), [1, 3], 50)
    assert C.shape == ref.shape and np.array_equal(C, ref)


def _geom_two_columns(M):
    return np.array([[20.0 * (m % 2), 20.0 * (m // 2)] for m in range(M)])


def test_detect_clips_and_features_processors(gpu, exe, tmp_path):
    M, N = 12, 300_000
    X = _recording(M, N, 6, units=8)
    Xw = oc.whiten_apply(X.astype(np.float32),
                         oc.whitening_matrix(oc.whiten_covariance(X.astype(np.float32))))
    mdaio.writemda(tmp_path / "pre.mda", Xw)
    geom = _geom_two_columns(M)
    np.savetxt(tmp_path / "geom.csv", geom, delimiter=",", fmt="%g")
    # detection: rows (central channel, time) in the ms3alg firings order
    run("msamd.detect_events", f"--timeseries={tmp_path / 'pre.mda'}",
        f"--event_times_out={tmp_path / 'ev.mda'}", "--detect_sign=-1", "--detect_threshold=3",
        "--detect_interval=10", "--_request_num_threads=8")
    ev = mdaio.readmda(tmp_path / "ev.mda")       # (L, 2)
    counts, times = oc.ms3alg_detect(Xw, 3, 10, -1, num_threads=8)
    ch_ref, t_ref = oc.firings_order(counts, times)
    say(f"  detect: {len(t_ref)} events")
    assert np.array_equal(ev[:, 0], ch_ref + 1) and np.array_equal(ev[:, 1], t_ref)
    # neighborhood clips of channel 3 (radius 30 um: 3-5 channels)
    run("msamd.extract_neighborhood_clips", f"--timeseries={tmp_path / 'pre.mda'}",
        f"--event_times={tmp_path / 'ev.mda'}", f"--geom={tmp_path / 'geom.csv'}",
        f"--clips_out={tmp_path / 'nc.mda'}", "--central_channel=3", "--adjacency_radius=30",
        "--clip_size=50")
    C = mdaio.readmda(tmp_path / "nc.mda")
    nb = [c - 1 for c in oc.channels_from_geom(geom, 3, 30)]
    tt = ev[ev[:, 0] == 3, 1].astype(np.int64)
    ref = oc.extract_clips(Xw, tt - oc.clip_tmid(50), nb, 50)
    assert C.shape == ref.shape and np.array_equal(C, ref)
    # two-stage PCA features per neighborhood, columns in event order
    run("msamd.pca_features", f"--timeseries={tmp_path / 'pre.mda'}",
        f"--event_times={tmp_path / 'ev.mda'}", f"--geom={tmp_path / 'geom.csv'}",
        f"--features_out={tmp_path / 'f.mda'}", "--adjacency_radius=30", "--clip_size=50")
    F = mdaio.readmda(tmp_path / "f.mda")         # (L, 10)
    assert F.shape == (len(t_ref), 10)
    worst = 0.0
    for m in range(1, M + 1):
        sel = ev[:, 0] == m
        if sel.sum() < 20:
            continue
        nbm = [c - 1 for c in oc.channels_from_geom(geom, m, 30)]
        clips = oc.extract_clips(Xw, ev[sel, 1].astype(np.int64) - oc.clip_tmid(50), nbm, 50)
        f_o = oc.stage2_features(oc.dimension_reduce_clips(clips, 10, 10000), 10, 10000)
        f64 = oc.stage2_features_fp64(oc.dimension_reduce_clips_fp64(clips, 10, 10000), 10, 10000)
        e = np.max(np.abs(F[sel] - f64)) / np.max(np.abs(f64))
        eo = np.max(np.abs(F[sel] - f_o)) / np.max(np.abs(f_o))
        oo = np.max(np.abs(f_o - f64)) / np.max(np.abs(f64))
        assert eo < PCA_STAGE2_BOUND + oo
        worst = max(worst, e)
    say(f"  pca features: all {F.shape[1]} features, max rel err against FP64 {worst:.2e}")
    assert worst < PCA_STAGE2_BOUND


def test_templates_and_fit_stage_processors(gpu, tmp_path):
    """ms3.compute_templates (every cluster, and a cluster subset in list
    order) and msamd.fit_stage (ms3alg's global templates, then the fit)
    against the oracle: templates bit-identical, kept firings identical."""
    from tests.test_gpu_fit import _sorted_events
    X, times, labels = _sorted_events(12, 240_000, seed=7, units=6)
    keep = labels > 0
    times, labels = times[keep], labels[keep]
    mdaio.writemda(tmp_path / "pre.mda", X)
    F = np.stack([np.ones(len(times)), times.astype(np.float64), labels.astype(np.float64)])
    mdaio.writemda(tmp_path / "firings.mda", F.T.copy())   # (L, 3) rows -> 3 x L .mda
    run("ms3.compute_templates", f"--timeseries={tmp_path / 'pre.mda'}",
        f"--firings={tmp_path / 'firings.mda'}", f"--templates_out={tmp_path / 't.mda'}",
        "--clip_size=50")
    Tg = mdaio.readmda(tmp_path / "t.mda")                  # (K, T, M)
    To, _ = oc.compute_templates(X, times, labels, 6, 50)
    assert Tg.shape == To.shape and np.array_equal(Tg, To)
    run("ms3.compute_templates", f"--timeseries={tmp_path / 'pre.mda'}",
        f"--firings={tmp_path / 'firings.mda'}", f"--templates_out={tmp_path / 't2.mda'}",
        "--clip_size=50", "--clusters=5,2")
    T2 = mdaio.readmda(tmp_path / "t2.mda")
    assert np.array_equal(T2, To[[4, 1]])
    run("msamd.fit_stage", f"--timeseries={tmp_path / 'pre.mda'}",
        f"--firings={tmp_path / 'firings.mda'}", f"--firings_out={tmp_path / 'f2.mda'}",
        "--clip_size=50", "--reference_num_threads=8")
    F2 = mdaio.readmda(tmp_path / "f2.mda")                 # (Lk, 3)
    gt, _ = oc.global_templates(X, times, labels, 6, 50)
    used = oc.fit_stage_computer(X, times, labels, gt, 8)
    say(f"  fit stage processor: {F2.shape[0]} of {len(times)} kept (oracle {len(used)})")
    assert np.array_equal(F2[:, 1].astype(np.int64), times[used])
    assert np.array_equal(F2[:, 2].astype(np.int64), labels[used])
