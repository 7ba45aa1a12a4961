"""Benchmark: channel-samples/s through filter -> mask -> whiten -> detect -> clips -> PCA.

Workload (single GPU): the 384-channel 4-column dense probe, 30 kHz, 300 s
(9,000,000 samples), int16 raw input at 0.195 uV/bit with 300 units and
artifact bursts on 1% of samples (mountainsort_amd.synth.config2, seed 2),
generated on the GPU and resident in HBM before timing.  One step is one
pass of the whole hot path over the recording:
  bandpass 300-6000 Hz (reference windows, 8-thread chunk plan)
  -> mask_out_artifacts (threshold 6, interval 2000; the decisions, with the
     masked copy fused into the covariance and apply reads)
  -> whiten (covariance, whitening matrix, apply)
  -> detect (threshold 3, sign -1, interval 10, 8-thread chunk plan)
  -> clips (T = 50, neighborhoods of radius 40 um)
  -> PCA features (10 per channel, then 10 per neighborhood).

Prints one JSON line (see the contract in the repository docs).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def neighborhoods(M, columns, pitch, radius):
    """get_channels_from_geom for every channel (p_mountainsort3.cpp:473-491), 0-based."""
    import numpy as np
    from mountainsort_amd.synth import probe_geometry
    geom = probe_geometry(M, columns, pitch)
    d = np.sqrt(((geom[:, None, :] - geom[None, :, :]) ** 2).sum(-1))
    nb = []
    for m in range(M):
        others = [m2 for m2 in range(M) if m2 != m and (radius < 0 or d[m, m2] <= radius)]
        nb.append([m] + others)
    return nb


class Pipeline:
    def __init__(self, dev, M, N, cfg, radius=40.0, num_threads=8, with_pca=True, fused=True):
        import numpy as np
        import torch
        from mountainsort_amd import frontend as fe
        self.fe = fe
        self.torch = torch
        self.dev = dev
        self.M, self.N = M, N
        self.num_threads = num_threads
        self.with_pca = with_pca
        # fused: the clip gather runs inside the stage-1 PCA kernels (no clip array)
        self.fused = fused and with_pca
        nb = neighborhoods(M, cfg.columns, cfg.pitch, radius)
        self.nb_channels = torch.tensor(np.concatenate(nb), dtype=torch.int32, device=dev)
        self.nb_offsets = torch.tensor(np.concatenate([[0], np.cumsum([len(c) for c in nb])]),
                                       dtype=torch.int64, device=dev)
        self.widths = [len(c) for c in nb]
        # preallocated stage outputs
        self.filt = torch.empty((N, M), dtype=torch.float32, device=dev)
        self.white = torch.empty((N, M), dtype=torch.float32, device=dev)
        self.stats = {}
        self.pca_tiles = None   # default tiles until tune_pca() has run

    def tune_pca(self):
        """Autotune the PCA tiles on the last run's events (the same shapes every step)."""
        import numpy as np
        if self.fused:
            self.pca_tiles, ms = self.fe.pca_autotune_timeseries(
                self.white, self.last_times, self.last_offsets, self.nb_channels,
                self.nb_offsets, 50)
            return self.pca_tiles, ms
        off = self.last_offsets.cpu().numpy().astype(np.int64)
        L = off[1:] - off[:-1]
        self.pca_tiles, ms = self.fe.pca_autotune(self.last_clips, L, np.array(self.widths), 50)
        return self.pca_tiles, ms

    def run(self, raw, timers=None):
        fe, torch = self.fe, self.torch
        ev = {}

        def mark(name):
            if timers is not None:
                e = torch.cuda.Event(enable_timing=True)
                e.record()
                ev[name] = e

        mark("start")
        fe.bandpass_filter(raw, 30000, 300, 6000, 1000, num_threads=self.num_threads, out=self.filt)
        mark("bandpass")
        # mask decisions only: the covariance and the apply read the filtered
        # data with masked rows as zero, so the masked copy is never written
        rows, _ = fe.mask_rows(self.filt, 6.0, 2000)
        self.rows = rows
        mark("mask")
        C = fe.whiten_covariance(self.filt, row_keep=rows)
        mark("covariance")
        W = fe.whitening_matrix(C)
        mark("whitening_matrix")
        self.C, self.W = C, W
        fe.whiten_apply(self.filt, W, out=self.white, row_keep=rows, xxt=C)
        mark("apply")
        offsets, times = fe.detect_events(self.white, 3.0, 10, -1, num_threads=self.num_threads)
        mark("detect")
        feats = None
        clips = None
        if self.fused:
            # clip gather fused into the stage-1 PCA kernels
            feats = fe.pca_features_from_timeseries(self.white, times, offsets, self.nb_channels,
                                                    self.nb_offsets, 50, tiles=self.pca_tiles)
            mark("clips_pca")
        else:
            clips, clip_offsets = fe.extract_clips_batched(self.white, times, offsets,
                                                           self.nb_channels, self.nb_offsets, 50)
            mark("clips")
            if self.with_pca:
                feats = fe.pca_features(clips, clip_offsets, offsets, self.nb_offsets, 50,
                                        tiles=self.pca_tiles)
                mark("pca")
        self.last_clips, self.last_offsets, self.last_times = clips, offsets, times
        if timers is not None:
            torch.cuda.synchronize()
            names = list(ev.keys())
            for a, b in zip(names[:-1], names[1:]):
                timers.setdefault(b, []).append(ev[a].elapsed_time(ev[b]))
        self.stats["events"] = int(times.numel())
        return feats


# Single-launch kernels whose per-launch time the stage events measure (the
# stage's ABI call launches this kernel plus at most a few-microsecond helper).
# The bound is the roof the kernel's algorithmic work meets first: the apply
# reads one float32 sample row and writes one (8 bytes per channel-sample)
# against 2M flops per channel-sample, 96 flop/B at M = 384, below the
# 312 flop/B ridge of dense fp16 MFMA over HBM, so it is HBM-bound; the
# covariance reads 4 bytes against 2M flops (192 flop/B) and is priced
# against MFMA; the power method against fp64 (11 matvecs of 2M^2 per
# component).  split_terms is the MFMA products the kernel issues per
# algorithmic product (split-FP16 / split-BF16 pieces).
ROOFLINE_KERNELS = {
    # the filter stage is one library call of three kernels per window batch
    # (225 batches at 384 ch x 300 s); its launch is the call
    "bandpass": {"kernel": "bp_stage_a_wide,bp_stage_b,bp_stage_c", "bound": "hbm",
                 "peak": 8000.0, "unit": "GB/s", "bytes_per_chs": 6},
    "apply": {"kernel": "apply_kernel", "bound": "hbm", "peak": 8000.0, "unit": "GB/s",
              "bytes_per_chs": 8, "mfma_peak": 2500.0, "dtype_peak": "fp16 dense",
              "split_terms": 3},
    "covariance": {"kernel": "cov_all_kernel", "bound": "mfma", "peak": 2500.0,
                   "unit": "TFLOP/s", "dtype_peak": "bf16 dense", "split_terms": 3},
    "whitening_matrix": {"kernel": "whitening_power_kernel", "bound": "mfma", "peak": 78.6,
                         "unit": "TFLOP/s", "dtype_peak": "fp64", "split_terms": 1},
}


def roofline_of(stage, ms, M, N, traffic_file):
    k = ROOFLINE_KERNELS[stage]
    if stage == "bandpass":
        alg = float(k["bytes_per_chs"]) * M * N
        achieved = alg / (ms * 1e-3) / 1e9
        out = {"bound": "hbm", "kernel": k["kernel"], "achieved": round(achieved, 1),
               "peak": k["peak"], "unit": k["unit"], "frac": round(achieved / k["peak"], 5),
               "traffic": None, "algorithmic_bytes_per_launch": alg, "launch_ms": round(ms, 4)}
        try:
            with open(traffic_file) as f:
                t = json.load(f)
            if t.get("kernel") == k["kernel"] and "traffic_bytes" in t:
                out["traffic"] = t["traffic_bytes"]
                out["traffic_source"] = os.path.basename(traffic_file)
        except (OSError, ValueError):
            pass
        return out
    if stage == "whitening_matrix":
        flops = 11.0 * 2.0 * M * M * M
    else:
        flops = 2.0 * M * M * N
    tflops = flops / (ms * 1e-3) / 1e12
    mfma_peak = k.get("mfma_peak", k["peak"])
    mfma = {"algorithmic_flops_per_launch": flops, "mfma_TFLOPs": round(tflops, 3),
            "peak_dtype": k["dtype_peak"], "mfma_terms_per_product": k["split_terms"],
            "mfma_issue_frac": round(tflops * k["split_terms"] / mfma_peak, 5)}
    if k["bound"] == "hbm":
        alg = float(k["bytes_per_chs"]) * M * N
        achieved = alg / (ms * 1e-3) / 1e9
        out = {"bound": "hbm", "kernel": k["kernel"], "achieved": round(achieved, 1),
               "peak": k["peak"], "unit": k["unit"], "frac": round(achieved / k["peak"], 5),
               "traffic": None, "algorithmic_bytes_per_launch": alg,
               "launch_ms": round(ms, 4), **mfma}
    else:
        out = {"bound": k["bound"], "kernel": k["kernel"], "achieved": round(tflops, 3),
               "peak": k["peak"], "unit": k["unit"], "frac": round(tflops / k["peak"], 5),
               "traffic": None, "launch_ms": round(ms, 4), **mfma}
    try:
        with open(traffic_file) as f:
            t = json.load(f)
        if t.get("kernel") == k["kernel"] and "traffic_bytes" in t:
            out["traffic"] = t["traffic_bytes"]
            out["traffic_source"] = os.path.basename(traffic_file)
    except (OSError, ValueError):
        pass
    return out


def learned_leg(dev, pipe, cfg, reps=3, train_seconds=60.0, steps=4000):
    """The learned detector: trained on a separate 60 s recording of the same
    probe (seed 4) labeled by plain detection, then run on the benchmark's
    whitened recording (seed 2) next to plain detection.  Reports recall and
    precision against plain detection (+-1 sample), window classes, and the
    detection and detect->clips->PCA times of both forms."""
    import numpy as np
    import torch
    from mountainsort_amd import frontend as fe
    from mountainsort_amd import learned, synth

    fe_ = fe
    Ntr = int(round(train_seconds * 30000))
    cfg_tr = synth.config2(N=Ntr, seed=4)
    raw_tr = synth.generate(cfg_tr, device=dev)
    tr = Pipeline(dev, cfg_tr.M, Ntr, cfg_tr, with_pca=False)
    tr.run(raw_tr)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    params, trace = learned.train_detector(tr.white, 3.0, 10, -1, num_threads=pipe.num_threads,
                                           steps=steps, seed=4)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    trace = trace.cpu().numpy()
    del tr, raw_tr
    torch.cuda.empty_cache()

    X = pipe.white
    det = learned.LearnedDetection(params)

    def plain_detect():
        return fe_.detect_events(X, 3.0, 10, -1, num_threads=pipe.num_threads)

    def learned_detect():
        off, t, _ = det(X, 3.0, 10, -1, num_threads=pipe.num_threads)
        return off, t

    def features(off, t):
        if pipe.fused:
            fe_.pca_features_from_timeseries(X, t, off, pipe.nb_channels, pipe.nb_offsets, 50,
                                             tiles=pipe.pca_tiles)
            return
        clips, co = fe_.extract_clips_batched(X, t, off, pipe.nb_channels, pipe.nb_offsets, 50)
        if pipe.with_pca:
            fe_.pca_features(clips, co, off, pipe.nb_offsets, 50, tiles=pipe.pca_tiles)

    def timed(fn, with_features):
        ms_d, ms_all = [], []
        for _ in range(reps + 1):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            c = torch.cuda.Event(enable_timing=True)
            a.record()
            off, t = fn()
            b.record()
            if with_features:
                features(off, t)
            c.record()
            torch.cuda.synchronize()
            ms_d.append(a.elapsed_time(b))
            ms_all.append(a.elapsed_time(c))
        return float(np.mean(ms_d[1:])), float(np.mean(ms_all[1:])), off, t

    pd, pall, off_p, t_p = timed(plain_detect, True)
    ld, lall, off_l, t_l = timed(learned_detect, True)
    _, _, counts = det(X, 3.0, 10, -1, num_threads=pipe.num_threads)
    hit_p, _, n_p = learned.match_events(off_l, t_l, off_p, t_p)
    hit_l, _, n_l = learned.match_events(off_p, t_p, off_l, t_l)
    out = {
        "training": {"recording": f"384 ch x {train_seconds:g} s (synth.config2 seed 4)",
                     "steps": steps, "batch": 256, "seg_windows": 16, "pos_weight": 20.0,
                     "seconds": round(train_s, 3),
                     "loss_first_last": [round(float(trace[:50].mean()), 4),
                                         round(float(trace[-50:].mean()), 4)]},
        "score_band": [0.2, 0.8],
        "recall_vs_plain": round(hit_p / max(n_p, 1), 5),
        "precision_vs_plain": round(hit_l / max(n_l, 1), 5),
        "events_plain": n_p, "events_learned": n_l,
        "windows_skipped_uncertain_present": counts,
        "detect_ms": {"plain": round(pd, 3), "learned": round(ld, 3)},
        "detect_clips_pca_ms": {"plain": round(pall, 3), "learned": round(lall, 3)},
        "speedup_detect": round(pd / ld, 3),
        "speedup_detect_clips_pca": round(pall / lall, 3),
    }
    log("[bench] learned: " + json.dumps(out))
    return out


def exact_whitening_leg(pipe, reps=3):
    """The exact whitening mode on the benchmark's covariance: W = C^{-1/2}
    (FP64 Newton-Schulz on the GPU) against the reference-method W of the
    timed steps.  Reports the deviation, both matrices' whitening error
    |W C W - I|_F / sqrt(M), iterations and time."""
    import numpy as np
    import torch
    from mountainsort_amd import frontend as fe
    C = pipe.C
    ms = []
    for _ in range(reps + 1):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        We, (iters, resid) = fe.whitening_matrix(C, exact=True, return_info=True)
        b.record()
        torch.cuda.synchronize()
        ms.append(a.elapsed_time(b))
    Cn = C.cpu().numpy()
    Wr = pipe.W.cpu().numpy()
    We = We.cpu().numpy()
    M = Cn.shape[0]
    I = np.eye(M)
    out = {
        "ms": round(float(np.median(ms[1:])), 3), "iterations": iters, "residual": resid,
        "deviation_from_reference_method": float(np.linalg.norm(We - Wr) / np.linalg.norm(Wr)),
        "whitening_error_exact": float(np.linalg.norm(We @ Cn @ We - I) / np.sqrt(M)),
        "whitening_error_reference_method": float(np.linalg.norm(Wr @ Cn @ Wr - I) / np.sqrt(M)),
    }
    log("[bench] exact whitening: " + json.dumps(out))
    return out


def pca_fusion_leg(pipe, reps=3):
    """Clips + PCA on the benchmark's events, unfused (clip array written by
    extract_clips_batched, then the batched PCA; tiles autotuned on the
    clips) against the fused form the timed steps run (clip gather inside
    the stage-1 PCA kernels; tiles autotuned on the timeseries).  Also
    checks that the two forms give bit-identical features at equal tiles."""
    import numpy as np
    import torch
    from mountainsort_amd import frontend as fe
    X, off, t = pipe.white, pipe.last_offsets, pipe.last_times
    clips, co = fe.extract_clips_batched(X, t, off, pipe.nb_channels, pipe.nb_offsets, 50)
    eo = off.cpu().numpy().astype(np.int64)
    tiles_u, _ = fe.pca_autotune(clips, eo[1:] - eo[:-1], np.array(pipe.widths), 50)
    del clips
    torch.cuda.empty_cache()

    def unfused():
        c, co_ = fe.extract_clips_batched(X, t, off, pipe.nb_channels, pipe.nb_offsets, 50)
        return fe.pca_features(c, co_, off, pipe.nb_offsets, 50, tiles=tiles_u)

    def fused(tiles):
        return fe.pca_features_from_timeseries(X, t, off, pipe.nb_channels, pipe.nb_offsets, 50,
                                               tiles=tiles)

    def timed(fn):
        ms = []
        for _ in range(reps + 1):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            torch.cuda.synchronize()
            ms.append(a.elapsed_time(b))
        return float(np.mean(ms[1:]))

    ms_u = timed(unfused)
    ms_f = timed(lambda: fused(pipe.pca_tiles))
    r1, _, f1, _ = unfused()
    r1, f1 = r1.clone(), f1.clone()
    r2, _, f2, _ = fused(tiles_u)
    same = bool(torch.equal(r1, r2) and torch.equal(f1, f2))
    out = {"events": int(t.numel()), "tiles_unfused": list(tiles_u),
           "tiles_fused": list(pipe.pca_tiles) if pipe.pca_tiles else None,
           "ms": {"unfused_clips_plus_pca": round(ms_u, 3), "fused": round(ms_f, 3)},
           "features_identical_at_equal_tiles": same}
    log("[bench] pca fusion: " + json.dumps(out))
    return out


def pca_batching_leg(dev, reps=3):
    """Batched PCA workload: 384 neighborhoods of 5-12 channels with
    log-uniform 100-200000 events (synth.pca_batch_shapes seed 3), 50-sample
    clips (synth.pca_batch_clips), 10 + 10 features.  The tiles come from the
    autotuner; the batched plan is timed as plain launches and as one HIP
    graph launch, against one plan per neighborhood launched in turn (same
    tiles, so the features are bit-identical)."""
    import numpy as np
    import torch
    from mountainsort_amd import frontend as fe
    from mountainsort_amd import synth
    L, W = synth.pca_batch_shapes(G=384, seed=3)
    clips, co = synth.pca_batch_clips(L, W, T=50, seed=3, device=dev)
    torch.cuda.synchronize()
    tiles, tune_ms = fe.pca_autotune(clips, L, W, 50)
    plan = fe.PcaPlan(L, W, 50, tiles=tiles, device=dev)
    dplan = fe.PcaPlan(L, W, 50, device=dev)
    singles = [fe.PcaPlan(L[g:g + 1], W[g:g + 1], 50, tiles=tiles, device=dev)
               for g in range(len(L))]
    views = [clips[co[g]:co[g + 1]] for g in range(len(L))]

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        out = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            torch.cuda.synchronize()
            out.append(a.elapsed_time(b))
        return float(np.median(out))

    ms_launch = timed(lambda: plan.run(clips))
    ms_graph = timed(lambda: plan.run(clips, graph=True))
    ms_default = timed(lambda: dplan.run(clips))

    def per_neighborhood():
        for g, p1 in enumerate(singles):
            if L[g]:
                p1.run(views[g])

    ms_single = timed(per_neighborhood)
    _, ro, feat, fo = plan.run(clips, graph=True)
    torch.cuda.synchronize()
    same = all(torch.equal(singles[g].features[:fo[g + 1] - fo[g]], feat[fo[g]:fo[g + 1]])
               for g in range(len(L)) if L[g])
    events = int(L.sum())
    clip_bytes = int((L * W).sum()) * 50 * 4
    # algorithmic bytes: clips read once, reduced written and read once, features written
    alg = clip_bytes + 2 * int((L * W).sum()) * 10 * 4 + events * 10 * 4
    out = {
        "workload": "384 neighborhoods, widths U{5..12}, events log-U(100, 200000) "
                    "(synth.pca_batch_shapes seed 3), T = 50, 10 + 10 features",
        "events": events, "clip_GB": round(clip_bytes / 1e9, 3),
        "tiles": {"gram_columns_per_item": tiles[0], "stage1_events_per_item": tiles[1],
                  "stage2_events_per_item": tiles[2], "autotune_best_ms": round(tune_ms, 3)},
        "ms": {"batched_graph": round(ms_graph, 3), "batched_launches": round(ms_launch, 3),
               "batched_default_tiles": round(ms_default, 3),
               "per_neighborhood_launches": round(ms_single, 3)},
        "speedup_batched_vs_per_neighborhood": round(ms_single / ms_graph, 3),
        "achieved_GBs_batched_graph": round(alg / (ms_graph * 1e-3) / 1e9, 1),
        "per_neighborhood_features_identical": bool(same),
    }
    for p1 in singles:
        p1.close()
    plan.close()
    dplan.close()
    del clips
    torch.cuda.empty_cache()
    log("[bench] pca batching: " + json.dumps(out))
    return out


# The port's stage times against the reference's own binaries, both at 8
# threads on 64 channels x 1.8 M samples (float32, 300-6000 Hz, threshold 3)
# in an 8-core 2.7 GHz Xeon container: (reference seconds, port seconds).
# The reference's detect+clips figure is its ms3alg read phase.
CPU_CALIBRATION = {
    "bandpass": (3.205, 4.70),
    "mask": (1.911, 0.503),
    "whiten": (7.278, 8.68),
    "detect_clips": (2.688, 2.32),
}


def _calibration_factor(stage):
    """Scale a port stage time down to the reference's speed where the port
    is the slower one (never up: the baseline errs towards a faster CPU)."""
    key = stage if stage in CPU_CALIBRATION else (
        "whiten" if stage in ("covariance", "whitening_matrix", "apply") else None)
    if key is None:
        return 1.0
    ref, port = CPU_CALIBRATION[key]
    return min(1.0, ref / port)


def cpu_threads():
    """Host threads for the baseline: this process's CPU set, at most 16 (the
    host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(M, N_full, counts_full, widths, thread_counts, log_fn=log, device="cpu"):
    """Time the reference's CPU path (oracle.cpu_path: the reference's loops
    with its OpenMP work split) on a bounded sample of the same workload, at
    each thread count, and scale each stage to the full recording by its own
    work plan.  Every timed unit is a whole unit of the reference's plan,
    one per thread, except the whitening chunks (1e4 of the reference's 1e5
    samples; the loop's cost per sample has no per-chunk term but an M x M
    merge):
      bandpass   T windows of chunk+120000 samples x 384 ch      x ceil(windows/T)
      mask       384 x 240000, one thread (as the reference)      x N/240000
      covariance T chunks of 1e4 samples                          x ceil(chunks/T) x 10
      W matrix   the 192 x 192 leading block, one thread          x (384/192)^4
      apply      T chunks of 1e4 samples                          x ceil(chunks/T) x 10
      detect+clips  the ms3alg read phase on 384 x 240000         x N/240000
      PCA        T whole neighborhoods (the full run's event counts and widths,
                 every 384/T-th channel), one batch                x ceil(384/T)
    The sample recording is generated on `device` (not timed).  Returns the
    cpu_baseline JSON object for the largest thread count, with the others
    under "by_threads"."""
    import math

    import numpy as np
    from mountainsort_amd import synth
    from oracle import core as oc
    from oracle import cpu_path as cp
    t_all = time.perf_counter()
    Ns = 240000
    cfg = synth.config2(N=Ns, seed=2)
    raw = synth.generate(cfg, device=device).cpu().numpy().astype(np.float32)
    T0 = max(thread_counts)
    filt = cp.bandpass_mt(raw, 30000, 300, 6000, num_threads=T0)
    nb = neighborhoods(M, cfg.columns, cfg.pitch, 40.0)
    shared = {}
    t = time.perf_counter()
    masked, _, _ = oc.mask_out_artifacts(filt, 6.0, 2000)
    shared["mask"] = (time.perf_counter() - t) * N_full / Ns
    Wb = min(M, 192)
    Cb = oc.whiten_covariance(np.ascontiguousarray(masked[:20000, :Wb]))
    t = time.perf_counter()
    oc.whitening_matrix(Cb)
    shared["whitening_matrix"] = (time.perf_counter() - t) * (M / Wb) ** 4
    C = oc.whiten_covariance(masked)
    W = oc.whitening_matrix(C) if M <= 96 else None
    if W is None:   # the same whitening from an FP64 eigendecomposition (untimed)
        w, V = np.linalg.eigh(oc.covariance_fp64(masked))
        W = (V / np.sqrt(np.maximum(w, 1e-12))) @ V.T
    white = oc.quantize_1e4((masked.astype(np.float64) @ W.T).astype(np.float32))
    log_fn(f"[bench] cpu baseline: sample prepared in {time.perf_counter() - t_all:.1f}s")
    n_win = math.ceil(N_full / max(oc.bandpass_chunk_size(M, T0), 1))
    n_chunks = math.ceil(N_full / 100000)
    nc = 10000
    counts_full = np.asarray(counts_full, np.int64)
    by_threads = {}
    for T in sorted(thread_counts, reverse=True):
        secs = dict(shared)
        cs = oc.bandpass_chunk_size(M, T)
        t = time.perf_counter()
        cp.bandpass_mt(raw, 30000, 300, 6000, num_threads=T, chunk_size=cs,
                       windows=[min(60000, Ns - 1)] * T)
        secs["bandpass"] = (time.perf_counter() - t) * math.ceil(N_full / cs / T)
        Xc = np.ascontiguousarray(masked[:T * nc])
        t = time.perf_counter()
        cp.whiten_covariance_mt(Xc, num_threads=T, chunk_size=nc)
        secs["covariance"] = (time.perf_counter() - t) * math.ceil(n_chunks / T) * 100000 / nc
        t = time.perf_counter()
        cp.whiten_apply_mt(Xc, W, num_threads=T, chunk_size=nc)
        secs["apply"] = (time.perf_counter() - t) * math.ceil(n_chunks / T) * 100000 / nc
        t = time.perf_counter()
        clips = cp.ms3alg_clips_mt(white, nb, 3.0, 10, -1, 50, num_threads=T)
        secs["detect_clips"] = (time.perf_counter() - t) * N_full / Ns
        rng = np.random.default_rng(0)
        pick = np.unique(np.round(np.linspace(0, M - 1, T)).astype(np.int64))
        sample = []
        for g in pick:
            Lg = int(counts_full[g]) if g < len(counts_full) else int(np.mean(counts_full))
            src = clips[g]
            if len(src) == 0:
                src = rng.normal(size=(1, 50, widths[g])).astype(np.float32)
            sample.append(np.ascontiguousarray(np.resize(src, (Lg, 50, widths[g]))))
        t = time.perf_counter()
        cp.pca_mt(sample, num_threads=T)
        secs["pca"] = (time.perf_counter() - t) * math.ceil(M / len(pick))
        cal = {k: v * _calibration_factor(k) for k, v in secs.items()}
        by_threads[T] = {
            "value": M * N_full / sum(cal.values()),
            "value_uncalibrated": M * N_full / sum(secs.values()),
            "stage_seconds": {k: round(v, 2) for k, v in cal.items()},
            "stage_seconds_uncalibrated": {k: round(v, 2) for k, v in secs.items()},
            "total_seconds": round(sum(cal.values()), 1),
        }
        log_fn(f"[bench] cpu baseline T={T}: " + json.dumps(by_threads[T]["stage_seconds"]) +
               f" total {by_threads[T]['total_seconds']}s")
    wall = time.perf_counter() - t_all
    top = by_threads[T0]
    return {
        "value": top["value"],
        "unit": "channel-samples/s",
        "cores": T0,
        "kind": "port",
        "sample": ("the reference's CPU path restated with its OpenMP work split "
                   "(oracle/cpu_path.py), on a 384 x 240000 sample of the same synthetic "
                   "recording; whole reference work units, one per thread (filter windows, "
                   "ms3alg time chunks, PCA neighborhoods at the full run's event counts), "
                   "whitening chunks of 1e4 samples, each stage scaled to 384 x "
                   f"{N_full} by its work plan; stage times multiplied by min(1, reference/"
                   "port) from an 8-thread calibration against the reference binaries"),
        "stage_seconds": top["stage_seconds"],
        "total_seconds": top["total_seconds"],
        "by_threads": {str(k): v for k, v in by_threads.items()},
        "calibration_reference_over_port": {k: round(r / p, 3)
                                            for k, (r, p) in CPU_CALIBRATION.items()},
        "sample_wall_seconds": round(wall, 1),
    }


def file_to_file_leg(dev, cfg_name, seconds, workdir, log_fn=log):
    """The processor chain file to file, as the MountainLab runner would call
    it: ms3.bandpass_filter -> ms3.mask_out_artifacts -> ms3.whiten ->
    msamd.detect_events -> msamd.pca_features, each a separate msamd.mp
    process reading and writing .mda files in workdir (every intermediate is
    a file, as between the reference's processors).  The synthetic recording
    is written first (not timed).  Returns seconds per processor and the
    whole chain's channel-samples/s."""
    import subprocess

    import numpy as np
    from mountainsort_amd import mdaio, synth
    exe = os.path.join(ROOT, "mountainsort_amd", "bin", "msamd.mp")
    N = int(round(seconds * 30000))
    cfg = synth.config1(N=N, seed=1) if cfg_name == "1" else synth.config2(N=N, seed=2)
    os.makedirs(workdir, exist_ok=True)
    raw_path = os.path.join(workdir, "raw.mda")
    t0 = time.time()
    raw = synth.generate(cfg, device=dev).cpu().numpy()
    mdaio.writemda(raw_path, raw)
    geom = os.path.join(workdir, "geom.csv")
    np.savetxt(geom, synth.probe_geometry(cfg.M, cfg.columns, cfg.pitch), delimiter=",")
    del raw
    log_fn(f"[bench] file-to-file: wrote {cfg.M} x {N} {cfg_name == '2' and 'int16' or 'float32'} "
           f"recording in {time.time() - t0:.1f}s")
    f = lambda name: os.path.join(workdir, name)
    steps = [
        ("ms3.bandpass_filter", [f"--timeseries={raw_path}", f"--timeseries_out={f('filt.mda')}",
                                 "--samplerate=30000", "--freq_min=300", "--freq_max=6000",
                                 "--_request_num_threads=8"]),
        ("ms3.mask_out_artifacts", [f"--timeseries={f('filt.mda')}",
                                    f"--timeseries_out={f('masked.mda')}", "--threshold=6",
                                    "--interval_size=2000"]),
        ("ms3.whiten", [f"--timeseries={f('masked.mda')}", f"--timeseries_out={f('white.mda')}"]),
        ("msamd.detect_events", [f"--timeseries={f('white.mda')}",
                                 f"--event_times_out={f('events.mda')}", "--detect_threshold=3",
                                 "--detect_sign=-1", "--detect_interval=10",
                                 "--_request_num_threads=8"]),
        ("msamd.pca_features", [f"--timeseries={f('white.mda')}", f"--event_times={f('events.mda')}",
                                f"--geom={geom}", "--adjacency_radius=40", "--clip_size=50",
                                f"--features_out={f('features.mda')}"]),
    ]
    secs = {}
    sizes_done = {}
    for name, args in steps:
        t = time.perf_counter()
        r = subprocess.run([exe, name, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True, timeout=900)
        secs[name] = round(time.perf_counter() - t, 3)
        if r.returncode != 0:
            raise RuntimeError(f"{name} failed: {r.stderr[-1000:]}")
        log_fn(f"[bench] file-to-file: {name} {secs[name]:.2f}s")
        # inputs no later step reads (keeps the disk footprint to three arrays)
        for done in {"ms3.mask_out_artifacts": ["filt.mda"], "ms3.whiten": ["masked.mda"]}.get(name, []):
            sizes_done[done] = os.path.getsize(f(done))
            os.remove(f(done))
    total = sum(secs.values())
    sizes = dict(sizes_done, **{"white.mda": os.path.getsize(f("white.mda"))})
    for n in ("raw.mda", "filt.mda", "masked.mda", "white.mda", "events.mda", "features.mda"):
        try:
            os.remove(f(n))
        except OSError:
            pass
    return {"workload": f"config {cfg_name}: {cfg.M} ch x {N} samples", "seconds": secs,
            "total_seconds": round(total, 3), "channel_samples_per_s": cfg.M * N / total,
            "intermediate_bytes": sizes, "workdir": workdir}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seconds", type=float, default=300.0, help="recording length")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-pca-batching", action="store_true",
                    help="leave out the batched-PCA leg (it runs after the timed steps)")
    ap.add_argument("--file-to-file", metavar="DIR",
                    help="instead of the device-resident step: time the processor chain "
                         ".mda file to file in DIR (config --f2f-config, --seconds long)")
    ap.add_argument("--f2f-config", choices=["1", "2"], default="2")
    ap.add_argument("--unfused", action="store_true",
                    help="time clips and PCA as two stages (clip array written, then read)")
    ap.add_argument("--pmc-pass", choices=sorted(ROOFLINE_KERNELS) + ["pca", "detect", "tcn"],
                    help="for rocprofv3 --pmc runs: one pipeline pass, then a 1 GiB "
                         "elementwise negation (counter calibration) and the named stage, "
                         "three times each")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the pipeline runs on one GPU only; sharding channel "
                                   "blocks and time spans over several GPUs is not "
                                   "implemented, so --gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    import torch
    from mountainsort_amd import synth

    dev = torch.device("cuda:0")
    if args.file_to_file:
        print(json.dumps({"file_to_file": file_to_file_leg(dev, args.f2f_config, args.seconds,
                                                           args.file_to_file)}), flush=True)
        return
    N = int(round(args.seconds * 30000))
    cfg = synth.config2(N=N, seed=2)
    t0 = time.time()
    raw = synth.generate(cfg, device=dev)
    torch.cuda.synchronize()
    log(f"[bench] generated {cfg.M} x {N} int16 in {time.time() - t0:.1f}s")
    pipe = Pipeline(dev, cfg.M, N, cfg, fused=not args.unfused)
    if args.pmc_pass:
        from mountainsort_amd import frontend as fe
        pipe.run(raw)
        copy_bytes = 1 << 30
        src = torch.ones(copy_bytes // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):  # calibration: reads and writes copy_bytes (torch's neg kernel)
            torch.neg(src, out=dst)
        for _ in range(3):
            if args.pmc_pass == "bandpass":
                fe.bandpass_filter(raw, 30000, 300, 6000, num_threads=8, out=pipe.filt)
            elif args.pmc_pass == "apply":
                fe.whiten_apply(pipe.filt, pipe.W, out=pipe.white, row_keep=pipe.rows, xxt=pipe.C)
            elif args.pmc_pass == "covariance":
                fe.whiten_covariance(pipe.filt, row_keep=pipe.rows)
            elif args.pmc_pass == "detect":
                fe.detect_events(pipe.white, 3.0, 10, -1, num_threads=8)
            elif args.pmc_pass == "tcn":
                # the network's window scores (untrained weights: the kernel's
                # work does not depend on their values)
                from mountainsort_amd import learned
                learned.score(pipe.white, learned.init_params(0, device=dev))
            elif args.pmc_pass == "pca":
                fe.pca_features_from_timeseries(pipe.white, pipe.last_times, pipe.last_offsets,
                                                pipe.nb_channels, pipe.nb_offsets, 50,
                                                tiles=(2560, 32, 32))
            else:
                fe.whitening_matrix(pipe.C)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_pass": args.pmc_pass, "copy_bytes": copy_bytes}), flush=True)
        return
    for i in range(args.warmup):
        pipe.run(raw)
        torch.cuda.synchronize()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
        if i == 0:
            tiles, tms = pipe.tune_pca()
            log(f"[bench] PCA tiles {tiles} ({tms:.2f} ms in the autotuner)")
    timers = {}
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    start_ev = torch.cuda.Event(enable_timing=True)
    end_ev = torch.cuda.Event(enable_timing=True)
    start_ev.record()
    for i in range(args.steps):
        pipe.run(raw, timers=timers)
        log(f"[bench] step {i + 1}/{args.steps} done")
    end_ev.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = wall * 1000.0 / args.steps
    chs = cfg.M * N
    value = chs * args.steps / wall
    stage_ms = {k: float(np.mean(v)) for k, v in timers.items()}
    log("[bench] stage ms: " + json.dumps({k: round(v, 3) for k, v in stage_ms.items()}))

    # algorithmic bytes per channel-sample of each stage (input once, output once);
    # the mask's copy is fused into the covariance and apply reads, so the mask
    # stage reads its input once (interval norms) and writes no timeseries
    alg_bytes = {"bandpass": 2 + 4, "mask": 4, "covariance": 4, "apply": 4 + 4,
                 "detect": 4 + 4}
    rl = {}
    for k, b in alg_bytes.items():
        if k in stage_ms:
            gbs = chs * b / (stage_ms[k] * 1e-3) / 1e9
            rl[k] = {"achieved_GBs": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}
    # the single-launch kernel with the largest launch time in the step
    dom = max((k for k in ROOFLINE_KERNELS if k in stage_ms), key=lambda k: stage_ms[k])
    tname = "bandpass" if dom == "bandpass" else ROOFLINE_KERNELS[dom]["kernel"]
    traffic_file = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                                f"{tname}_384ch_300s_traffic.json")
    roofline = roofline_of(dom, stage_ms[dom], cfg.M, N, traffic_file)
    # the same figures for every stage that has one
    rooflines = {}
    for st_name in ROOFLINE_KERNELS:
        if st_name in stage_ms:
            tn = "bandpass" if st_name == "bandpass" else ROOFLINE_KERNELS[st_name]["kernel"]
            rooflines[st_name] = roofline_of(
                st_name, stage_ms[st_name], cfg.M, N,
                os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                             f"{tn}_384ch_300s_traffic.json"))
    total_alg = sum(alg_bytes.values()) * chs
    line = {
        "metric": "channel-samples/sec through filter->whiten->detect->clips->PCA; % HBM peak",
        "value": value,
        "unit": "channel-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"384-ch 4-column probe, 30 kHz, {args.seconds:g} s, int16 "
                               "0.195 uV/bit, 300 units, 1% artifacts (synth.config2 seed 2)",
                   "channels": cfg.M, "samples": N, "reference_threads": 8,
                   "adjacency_radius_um": 40, "clip_size": 50, "events": pipe.stats.get("events"),
                   "pca": pipe.with_pca, "pca_tiles": pipe.pca_tiles},
        "hbm_frac_pipeline": round(total_alg / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "stage_ms": {k: round(v, 3) for k, v in stage_ms.items()},
        "stage_roofline": rl,
        "roofline": roofline,
        "rooflines": rooflines,
        "cpu_baseline": None,
    }
    if not args.no_learned:
        line["learned"] = learned_leg(dev, pipe, cfg)
    line["whitening_exact"] = exact_whitening_leg(pipe)
    if pipe.fused:
        line["pca_fusion"] = pca_fusion_leg(pipe)
    if not args.no_pca_batching:
        line["pca_batching"] = pca_batching_leg(dev)
    if not args.no_cpu_baseline:
        off = pipe.last_offsets.cpu().numpy().astype(np.int64)
        T = cpu_threads()
        line["cpu_baseline"] = cpu_baseline(cfg.M, N, off[1:] - off[:-1], pipe.widths,
                                            sorted({T, min(T, 8)}), device=dev)
        line["speedup_vs_cpu"] = round(value / line["cpu_baseline"]["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
