"""Writes tests/golden/front_end_small.npz: expected outputs of the oracle
(oracle/, the CPU restatement of the reference) on small seeded inputs.

The reference ships no golden vectors for this path, so these fixtures are
regression anchors produced by the oracle, whose own correctness is pinned by
tests/test_oracle.py (closed forms, brute force, invariants).  Inputs are not
stored: they are regenerated from the seeds below with numpy's PCG64, whose
streams are fixed across numpy versions; `inputs()` is shared with the tests.
Large outputs are stored at fixed sample positions only.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PATH = os.path.join(HERE, "front_end_small.npz")


def _spikes(rng, N, M, n, amp):
    X = rng.normal(size=(N, M)) * 10.0
    t = np.arange(50)
    for _ in range(n):
        st = rng.integers(0, N - 50)
        w = -rng.uniform(*amp) * np.exp(-0.5 * ((t - 20) / 2.5) ** 2) \
            + 0.3 * rng.uniform(*amp) * np.exp(-0.5 * ((t - 30) / 5) ** 2)
        X[st:st + 50] += np.outer(w, rng.uniform(0.1, 1.0, M))
    return X


def inputs():
    """The seeded inputs of every fixture."""
    rng = np.random.default_rng(20240611)
    raw = np.clip(_spikes(rng, 130_000, 2, 40, (60, 200)), -32768, 32767).astype(np.int16)
    rec = _spikes(rng, 40_000, 8, 60, (60, 200)).astype(np.float32)
    rec[20_000:21_200] += 2500.0
    white = (rng.normal(size=(60_000, 5)) + 0.0).astype(np.float32)
    for _ in range(80):
        st = rng.integers(0, 60_000 - 50)
        white[st + 18:st + 24, rng.integers(0, 5)] -= rng.uniform(3, 8)
    clips = (rng.normal(size=(300, 50, 5)) +
             np.outer(np.sin(np.arange(50) / 4.0), np.linspace(0.2, 1.0, 5))[None]
             * rng.uniform(1, 6, 300)[:, None, None]).astype(np.float32)
    return {"raw": raw, "rec": rec, "white": white, "clips": clips}


def expected(inp):
    from oracle import core as oc
    out = {}
    # bandpass (2 channels, int16 in): 240000-point windows (94-thread chunk plan)
    bp = oc.bandpass_filter(inp["raw"], 30000, 300, 6000, num_threads=94)
    idx = np.random.default_rng(1).choice(bp.size, 4000, replace=False)
    out["bandpass_idx"] = idx.astype(np.int64)
    out["bandpass_val"] = bp.ravel()[idx]
    # artifact masking (bit-exact)
    Y, _, use_it = oc.mask_out_artifacts(inp["rec"], 6.0, 2000)
    out["mask_use_it"] = use_it
    out["mask_checksum"] = np.array([np.float64(Y).sum(), np.abs(np.float64(Y)).sum()])
    # whitening of the masked recording
    C = oc.whiten_covariance(Y)
    W = oc.whitening_matrix(C)
    out["whiten_cov"] = C
    out["whiten_W"] = W
    Yw = oc.whiten_apply(Y, W)
    out["whiten_rows"] = Yw[::997].copy()
    # detection (events of every channel, 8-thread chunk plan)
    cnt, times = oc.ms3alg_detect(inp["white"], 3.0, 10, -1, num_threads=8, chunk_size=20_000)
    out["detect_counts"] = cnt
    out["detect_times"] = times
    # clips at the detected events of channel 0 over channels [0, 2, 4]
    t0 = times[:cnt[0]] - oc.clip_tmid(50)
    out["clips_ch0"] = oc.extract_clips(inp["white"], t0, [0, 2, 4], 50)
    # two-stage PCA features of one neighborhood
    red = oc.dimension_reduce_clips(inp["clips"], 10, 10000)
    out["pca_reduced"] = red
    out["pca_features"] = oc.stage2_features(red, 10, 10000)
    return out


if __name__ == "__main__":
    np.savez_compressed(PATH, **expected(inputs()))
    print("wrote", PATH, os.path.getsize(PATH), "bytes")
