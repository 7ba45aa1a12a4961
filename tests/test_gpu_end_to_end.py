"""End-to-end GPU parity on a slice of the 64-channel probe recording:
bandpass -> whitening -> detection on the GPU against the oracle's chain
(FP64 filter, the reference's float whitening loop, its detection rule).

Event times must agree >= 99.9% within +-1 sample when the GPU applies the
oracle's whitening matrix; with the GPU's own matrix the agreement is reported
and held to the level the reference reaches against itself (its W is chaotic
in rounding at this size)."""
import numpy as np
import pytest
import torch

from oracle import core as oc
from tests.conftest import say

pytestmark = pytest.mark.gpu


def _agreement(counts_a, times_a, counts_b, times_b, tol=1):
    """Fraction of b's events with an event of a on the same channel within tol."""
    oa = np.concatenate([[0], np.cumsum(counts_a)])
    ob = np.concatenate([[0], np.cumsum(counts_b)])
    hit = 0
    for c in range(len(counts_b)):
        a = np.sort(times_a[oa[c]:oa[c + 1]])
        b = times_b[ob[c]:ob[c + 1]]
        if a.size == 0 or b.size == 0:
            continue
        i = np.clip(np.searchsorted(a, b), 1, a.size) - 1
        j = np.clip(i + 1, 0, a.size - 1)
        d = np.minimum(np.abs(a[i] - b), np.abs(a[j] - b))
        hit += int(np.sum(d <= tol))
    return hit / max(int(np.sum(counts_b)), 1)


def test_front_end_events_agree_with_the_oracle_chain(gpu):
    from mountainsort_amd import frontend as fe
    from mountainsort_amd import synth
    N = 600_000   # 20 s of the 64-channel 2-column probe (pink + white noise, 40 units)
    cfg = synth.config1(N=N, seed=1)
    raw = synth.generate(cfg, device="cpu").numpy()
    say(f"end to end: 64 ch x {N} samples")
    # oracle chain (the reference's arithmetic)
    filt_o = oc.bandpass_filter(raw, 30000, 300, 6000, 1000, num_threads=8).astype(np.float32)
    C_o = oc.whiten_covariance(filt_o)
    W_o = oc.whitening_matrix(C_o)
    white_o = oc.whiten_apply(filt_o, W_o)
    cnt_o, t_o = oc.ms3alg_detect(white_o, 3.0, 10, -1, num_threads=8)
    # GPU chain: filter, then apply with the oracle's W and with its own
    Xg = torch.from_numpy(raw).to(gpu)
    filt = fe.bandpass_filter(Xg, 30000, 300, 6000, 1000, num_threads=8)
    ferr = float(torch.max(torch.abs(filt.cpu() - torch.from_numpy(filt_o)))) / float(np.sqrt(np.mean(filt_o.astype(np.float64) ** 2)))
    say(f"  filtered: max |diff| / rms = {ferr:.2e}")
    assert ferr < 1e-4
    white_a = fe.whiten_apply(filt, torch.from_numpy(W_o).to(gpu))
    off_a, t_a = fe.detect_events(white_a, 3.0, 10, -1, num_threads=8)
    cnt_a = np.diff(off_a.cpu().numpy())
    t_a = t_a.cpu().numpy()
    agree = min(_agreement(cnt_a, t_a, cnt_o, t_o), _agreement(cnt_o, t_o, cnt_a, t_a))
    say(f"  events (oracle W on the GPU): {t_a.size} vs {t_o.size}, agreement {agree:.5f}")
    assert agree >= 0.999
    C = fe.whiten_covariance(filt)
    W = fe.whitening_matrix(C)
    white_g = fe.whiten_apply(filt, W, xxt=C)
    off_g, t_g = fe.detect_events(white_g, 3.0, 10, -1, num_threads=8)
    cnt_g = np.diff(off_g.cpu().numpy())
    t_g = t_g.cpu().numpy()
    agree_g = min(_agreement(cnt_g, t_g, cnt_o, t_o), _agreement(cnt_o, t_o, cnt_g, t_g))
    say(f"  events (GPU W): {t_g.size} vs {t_o.size}, agreement {agree_g:.5f}")
    assert agree_g >= 0.995
