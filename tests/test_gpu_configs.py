"""GPU parity on each benchmark workload at its own shape.

  config 0: the 4-channel, 60 s white-noise recording (synth.config0, seed 0)
            at full size, filter -> whiten -> detect -> clips -> PCA, against
            the oracle's chain;
  config 2: a 6 s slice of the 384-channel int16 probe with artifact bursts
            (synth.config2, seed 2), filter -> mask -> covariance/apply ->
            detect -> clips -> PCA;
  config 3: the whole 384-neighborhood batched PCA (synth.pca_batch_shapes
            seed 3), eight neighborhoods spanning the event counts and widths
            checked against the oracle;
  config 4: the learned detector trained on 60 s of the 384-channel probe
            (seed 4) and evaluated on a fresh 300 s recording (seed 5).

Bars: filtered samples within 1e-4 of the channel RMS of the FP64 oracle;
covariance within 2e-5 of the FP64 diagonal; whitened samples within one
1e-4 quantum of FP64 given the same W; event times and clips bit-exact given
the oracle's whitened data; end-to-end events >= 99.9% within +-1 sample;
every PCA component and feature within the conftest bounds."""
import numpy as np
import pytest
import torch

from oracle import core as oc
from tests.conftest import assert_pca_parity, say
from tests.test_gpu_end_to_end import _agreement

pytestmark = pytest.mark.gpu


def _filter_error(filt_gpu, filt_o):
    """max |gpu - oracle| / RMS, channel by channel, worst channel."""
    d = np.abs(filt_gpu.astype(np.float64) - filt_o).max(axis=0)
    rms = np.sqrt(np.mean(filt_o ** 2, axis=0))
    return float(np.max(d / rms))


def _check_clips_and_pca(gpu, white_o, cnt_o, t_o, nb, label):
    """Batched clip gather at the oracle's events on the oracle's whitened
    data (bit-exact) and batched PCA of those clips against the oracle."""
    from mountainsort_amd import frontend as fe
    T = 50
    eo = np.concatenate([[0], np.cumsum(cnt_o)]).astype(np.int64)
    no = np.concatenate([[0], np.cumsum([len(c) for c in nb])]).astype(np.int64)
    Xg = torch.from_numpy(white_o).to(gpu)
    times = torch.from_numpy(np.asarray(t_o, np.int64)).to(gpu)
    flat, co = fe.extract_clips_batched(Xg, times, eo, np.concatenate(nb).astype(np.int32), no, T)
    co = co.cpu().numpy()
    host = flat.cpu().numpy()
    tmid = oc.clip_tmid(T)
    for g in range(len(nb)):
        ref = oc.extract_clips(white_o, t_o[eo[g]:eo[g + 1]] - tmid, nb[g], T)
        got = host[co[g]:co[g + 1]].reshape(ref.shape)
        assert np.array_equal(got, ref), f"{label}: clips of neighborhood {g}"
    say(f"  {label}: clips bit-exact for {len(nb)} neighborhoods ({int(eo[-1])} events)")
    red, ro, feat, fo = fe.pca_features(flat, co, eo, no, T)
    red = red.cpu().numpy()
    feat = feat.cpu().numpy()
    worst1 = worst2 = 0.0
    for g in range(len(nb)):
        L, W = int(cnt_o[g]), len(nb[g])
        if L == 0:
            continue
        c = host[co[g]:co[g + 1]].reshape(L, T, W)
        r_o = oc.dimension_reduce_clips(c, 10, 10000)
        f_o = oc.stage2_features(r_o, 10, 10000)
        e1, e2 = assert_pca_parity(red[ro[g]:ro[g + 1]].reshape(L, 10, W), r_o,
                                   feat[fo[g]:fo[g + 1]].reshape(L, 10), f_o,
                                   f"{label} nbhd {g} L={L} W={W}", clips=c)
        worst1, worst2 = max(worst1, e1), max(worst2, e2)
    say(f"  {label}: PCA worst stage1 {worst1:.2e}, stage2 {worst2:.2e}")


def test_config0_full_chain(gpu):
    """4 channels x 1.8 M samples at the reference's 8-thread plan: the filter
    window is n = 1398101 (chirp-z form), neighborhoods are all 4 channels
    (adjacency_radius -1, no geometry)."""
    from mountainsort_amd import frontend as fe
    from mountainsort_amd import synth
    cfg = synth.config0()
    raw = synth.generate(cfg, device="cpu").numpy()
    N, M = raw.shape
    say(f"config 0: {M} ch x {N} samples, float32, seed {cfg.seed}")
    filt_o = oc.bandpass_filter(raw, 30000, 300, 6000, 1000, num_threads=8)
    f32 = filt_o.astype(np.float32)
    C_o = oc.whiten_covariance(f32)
    W_o = oc.whitening_matrix(C_o)
    white_o = oc.whiten_apply(f32, W_o)
    cnt_o, t_o = oc.ms3alg_detect(white_o, 3.0, 10, -1, num_threads=8)

    filt = fe.bandpass_filter(torch.from_numpy(raw).to(gpu), 30000, 300, 6000, 1000,
                              num_threads=8)
    ferr = _filter_error(filt.cpu().numpy(), filt_o)
    say(f"  filter: max |diff| / rms = {ferr:.2e}")
    assert ferr < 1e-4
    C = fe.whiten_covariance(filt)
    R = oc.covariance_fp64(f32)
    cerr = np.max(np.abs(C.cpu().numpy() - R)) / np.max(np.diag(R))
    W = fe.whitening_matrix(C)
    werr = np.max(np.abs(W.cpu().numpy() - W_o)) / np.max(np.abs(W_o))
    say(f"  covariance {cerr:.2e} of the diagonal; W vs oracle W {werr:.2e}")
    assert cerr < 2e-5
    assert werr < 1e-4
    # the apply given the oracle's filtered data and W: within one quantum
    Ya = fe.whiten_apply(torch.from_numpy(f32).to(gpu), torch.from_numpy(W_o).to(gpu)).cpu().numpy()
    dq = np.abs(np.round(Ya.astype(np.float64) / 1e-4) - np.round(white_o.astype(np.float64) / 1e-4))
    say(f"  apply: {np.mean(dq > 0):.5f} of samples one quantum off, max {dq.max():.0f}")
    assert dq.max() <= 1
    # detection on the oracle's whitened data: bit-exact
    off, tg = fe.detect_events(torch.from_numpy(white_o).to(gpu), 3.0, 10, -1, num_threads=8)
    assert np.array_equal(np.diff(off.cpu().numpy()), cnt_o)
    assert np.array_equal(tg.cpu().numpy(), t_o)
    # end to end: the GPU chain with its own W against the oracle chain
    white = fe.whiten_apply(filt, W, xxt=C)
    off_g, t_g = fe.detect_events(white, 3.0, 10, -1, num_threads=8)
    cnt_g, t_g = np.diff(off_g.cpu().numpy()), t_g.cpu().numpy()
    agree = min(_agreement(cnt_g, t_g, cnt_o, t_o), _agreement(cnt_o, t_o, cnt_g, t_g))
    say(f"  events: gpu chain {t_g.size}, oracle chain {t_o.size}, agreement {agree:.5f}")
    assert agree >= 0.999
    nb = [[m] + [c for c in range(M) if c != m] for m in range(M)]
    _check_clips_and_pca(gpu, white_o, cnt_o, t_o, nb, "config 0")


def test_config2_slice_chain(gpu):
    """384 channels x 6 s of the int16 probe with +-2000 artifact bursts: two
    filter windows of 240000, the mask before whitening, and the batched
    clips and PCA in 40 um neighborhoods."""
    from mountainsort_amd import frontend as fe
    from mountainsort_amd import synth
    N = 180_000
    cfg = synth.config2(N=N, seed=2)
    raw = synth.generate(cfg, device="cpu").numpy()
    M = raw.shape[1]
    say(f"config 2 slice: {M} ch x {N} samples, int16, seed {cfg.seed}")
    assert raw.dtype == np.int16
    filt_o = oc.bandpass_filter(raw, 30000, 300, 6000, 1000, num_threads=8)
    f32 = filt_o.astype(np.float32)
    masked_o, _, use_o = oc.mask_out_artifacts(f32, 6.0, 2000)
    say(f"  oracle mask keeps {int(use_o[:N // 2000].sum())} of {N // 2000} intervals")
    assert use_o[:N // 2000].sum() < N // 2000   # the bursts are masked

    filt = fe.bandpass_filter(torch.from_numpy(raw).to(gpu), 30000, 300, 6000, 1000,
                              num_threads=8)
    ferr = _filter_error(filt.cpu().numpy(), filt_o)
    say(f"  filter (int16 in): max |diff| / rms = {ferr:.2e}")
    assert ferr < 1e-4
    f32g = torch.from_numpy(f32).to(gpu)
    rows_o, use_g = fe.mask_rows(f32g, 6.0, 2000)
    assert np.array_equal(use_g.cpu().numpy(), use_o)
    rows, use_c = fe.mask_rows(filt, 6.0, 2000)
    assert np.array_equal(use_c.cpu().numpy(), use_o)
    C = fe.whiten_covariance(f32g, row_keep=rows_o)
    R = oc.covariance_fp64(masked_o)
    cerr = np.max(np.abs(C.cpu().numpy() - R)) / np.max(np.diag(R))
    say(f"  masked covariance: {cerr:.2e} of the diagonal")
    assert cerr < 2e-5
    W = fe.whitening_matrix(C)
    Wn = W.cpu().numpy()
    # W = C diag(sigma^-1/2) C^T summed per entry: symmetric up to rounding
    assert np.max(np.abs(Wn - Wn.T)) <= 1e-12 * np.max(np.abs(Wn))
    q = np.linalg.norm(Wn @ R @ Wn - np.eye(M)) / np.sqrt(M)
    say(f"  W (reference method): |W C W - I|_F / sqrt(M) = {q:.3f}")
    assert q < 0.1
    Y = fe.whiten_apply(f32g, W, row_keep=rows_o, xxt=C).cpu().numpy()
    k64 = np.floor((masked_o.astype(np.float64) @ Wn.T) / 1e-4 + 0.5)
    d = np.abs(np.round(Y.astype(np.float64) / 1e-4) - k64)
    say(f"  apply: {np.mean(d > 0):.5f} of samples one quantum off, max {d.max():.0f}")
    assert d.max() <= 1
    white_o = (k64 * 1e-4).astype(np.float32)
    del Y, k64, d
    cnt_o, t_o = oc.ms3alg_detect(white_o, 3.0, 10, -1, num_threads=8)
    off, tg = fe.detect_events(torch.from_numpy(white_o).to(gpu), 3.0, 10, -1, num_threads=8)
    assert np.array_equal(np.diff(off.cpu().numpy()), cnt_o)
    assert np.array_equal(tg.cpu().numpy(), t_o)
    say(f"  detection bit-exact: {t_o.size} events")
    # end to end: GPU chain (own filter, mask, covariance, W, apply) against the
    # oracle chain given the GPU chain's W
    Cc = fe.whiten_covariance(filt, row_keep=rows)
    Wc = fe.whitening_matrix(Cc)
    white = fe.whiten_apply(filt, Wc, row_keep=rows, xxt=Cc)
    off_g, t_g = fe.detect_events(white, 3.0, 10, -1, num_threads=8)
    cnt_g, t_g = np.diff(off_g.cpu().numpy()), t_g.cpu().numpy()
    wo2 = oc.quantize_1e4((masked_o.astype(np.float64) @ Wc.cpu().numpy().T).astype(np.float32))
    cnt_o2, t_o2 = oc.ms3alg_detect(wo2, 3.0, 10, -1, num_threads=8)
    agree = min(_agreement(cnt_g, t_g, cnt_o2, t_o2), _agreement(cnt_o2, t_o2, cnt_g, t_g))
    say(f"  events: gpu chain {t_g.size}, oracle chain {t_o2.size}, agreement {agree:.5f}")
    assert agree >= 0.999
    del white, wo2
    geom = synth.probe_geometry(M, cfg.columns, cfg.pitch)
    nb = [[c - 1 for c in oc.channels_from_geom(geom, m + 1, 40.0)] for m in range(M)]
    _check_clips_and_pca(gpu, white_o, cnt_o, t_o, nb, "config 2")


def test_config3_batch_against_oracle(gpu):
    """The full batching workload (384 neighborhoods, widths 5-12, events
    log-uniform 100-200000, about 9.4 M clips) with default and autotuned
    tiles; eight neighborhoods that span the counts and widths are checked
    against the oracle, and the batched features equal one plan per
    neighborhood bit for bit."""
    from mountainsort_amd import frontend as fe
    from mountainsort_amd import synth
    L, W = synth.pca_batch_shapes(G=384, seed=3)
    say(f"config 3: {len(L)} neighborhoods, {int(L.sum())} events, widths {W.min()}-{W.max()}")
    clips, co = synth.pca_batch_clips(L, W, T=50, seed=3, device=gpu)
    eo = np.concatenate([[0], np.cumsum(L)])
    no = np.concatenate([[0], np.cumsum(W)])
    red, ro, feat, fo = fe.pca_features(clips, co, eo, no, 50)
    red, feat = red.cpu().numpy(), feat.cpu().numpy()
    tiles, ms = fe.pca_autotune(clips, L, W, 50)
    say(f"  autotuned tiles {tiles} ({ms:.2f} ms)")
    plan = fe.PcaPlan(L, W, 50, tiles=tiles, device=gpu)
    _, ro2, feat2, fo2 = plan.run(clips, graph=True)
    torch.cuda.synchronize()
    red2 = plan.reduced.cpu().numpy()
    feat2 = feat2.cpu().numpy()
    order = np.argsort(L)
    pick = sorted({int(order[0]), int(order[-1]), int(np.argmin(W)), int(np.argmax(W)),
                   *(int(order[int(q * (len(L) - 1))]) for q in (0.25, 0.5, 0.75, 0.9))})
    for g in pick:
        Lg, Wg = int(L[g]), int(W[g])
        host = clips[co[g]:co[g + 1]].cpu().numpy().reshape(Lg, 50, Wg)
        r_o = oc.dimension_reduce_clips(host, 10, 10000)
        f_o = oc.stage2_features(r_o, 10, 10000)
        r64 = oc.dimension_reduce_clips_fp64(host, 10, 10000)
        f64 = oc.stage2_features_fp64(r64, 10, 10000)
        assert_pca_parity(red[ro[g]:ro[g + 1]].reshape(Lg, 10, Wg), r_o,
                          feat[fo[g]:fo[g + 1]].reshape(Lg, 10), f_o,
                          f"g={g} L={Lg} W={Wg} default tiles", fp64=(r64, f64))
        assert_pca_parity(red2[ro2[g]:ro2[g + 1]].reshape(Lg, 10, Wg), r_o,
                          feat2[fo2[g]:fo2[g + 1]].reshape(Lg, 10), f_o,
                          f"g={g} L={Lg} W={Wg} tiles {tiles}", fp64=(r64, f64))
        p1 = fe.PcaPlan(L[g:g + 1], W[g:g + 1], 50, tiles=tiles, device=gpu)
        r1, _, f1, _ = p1.run(clips[co[g]:co[g + 1]].contiguous())
        assert np.array_equal(r1.cpu().numpy()[:ro2[g + 1] - ro2[g]], red2[ro2[g]:ro2[g + 1]])
        assert np.array_equal(f1.cpu().numpy()[:fo2[g + 1] - fo2[g]], feat2[fo2[g]:fo2[g + 1]])
        p1.close()
    plan.close()


def _whitened_config2(gpu, N, seed):
    """The benchmark chain on synth.config2(N, seed): filter, mask decisions,
    masked covariance, reference-method W, apply."""
    from mountainsort_amd import frontend as fe
    from mountainsort_amd import synth
    raw = synth.generate(synth.config2(N=N, seed=seed), device=gpu)
    filt = fe.bandpass_filter(raw, 30000, 300, 6000, 1000, num_threads=8)
    del raw
    rows, _ = fe.mask_rows(filt, 6.0, 2000)
    C = fe.whiten_covariance(filt, row_keep=rows)
    W = fe.whitening_matrix(C)
    white = fe.whiten_apply(filt, W, row_keep=rows, xxt=C)
    del filt
    torch.cuda.empty_cache()
    return white


def test_config4_learned_384(gpu):
    """Learned detector at the 384-channel shape: trained on 60 s (seed 4)
    labeled by plain detection, evaluated on a fresh 300 s recording
    (seed 5) with plain detection deciding the 0.2-0.8 band."""
    from mountainsort_amd import frontend as fe
    from mountainsort_amd import learned
    say("config 4: training on 384 ch x 60 s (seed 4)")
    Xtr = _whitened_config2(gpu, 1_800_000, 4)
    params, trace = learned.train_detector(Xtr, 3.0, 10, -1, num_threads=8, steps=4000, seed=4)
    trace = trace.cpu().numpy()
    say(f"  loss first 50 {trace[:50].mean():.4f}, last 50 {trace[-50:].mean():.4f}")
    del Xtr
    torch.cuda.empty_cache()
    say("config 4: evaluation on 384 ch x 300 s (seed 5)")
    Xte = _whitened_config2(gpu, 9_000_000, 5)
    off_p, t_p = fe.detect_events(Xte, 3.0, 10, -1, num_threads=8)
    off_l, t_l, counts = learned.LearnedDetection(params)(Xte, 3.0, 10, -1, num_threads=8)
    hit_p, _, n_p = learned.match_events(off_l, t_l, off_p, t_p)
    hit_l, _, n_l = learned.match_events(off_p, t_p, off_l, t_l)
    recall, precision = hit_p / max(n_p, 1), hit_l / max(n_l, 1)
    say(f"  recall {recall:.4f} precision {precision:.4f} ({n_l} vs {n_p} plain events), "
        f"windows skipped/uncertain/present {counts}")
    assert n_p > 100_000
    assert recall >= 0.99
    assert precision >= 0.99
