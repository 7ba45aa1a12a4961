"""Learned spike-presence detector on device-resident tensors (C ABI: msamd_tcn_*).

The network scores 32-sample windows of every channel (see include/msamd.h).
Detection then runs only in the windows it proposes (score >= score_lo),
where the reference's detection rule (detect_events.cpp:69-135 over the
ms3alg time-chunk plan, p_mountainsort3.cpp:170-216) decides every sample.  Training labels
come from plain detection (frontend.detect_events) on a training recording;
training runs on the GPU through msamd_tcn_train.
"""
import ctypes
import math

import torch

from . import _capi
from .frontend import (DETECT_PADDING, _check_ts, _ptr, _stream, _workspace, detect_chunk_size,
                       detect_events)

WINDOW = 32


def num_params():
    return int(_capi.lib().msamd_tcn_num_params())


def init_params(seed=0, device="cuda"):
    p = torch.empty(num_params(), dtype=torch.float32, device=device)
    _capi.call("msamd_tcn_init", _ptr(p), ctypes.c_uint64(seed), _stream(p.device))
    return p


def labels_from_events(offsets, times, M, N):
    """(labels (nW, M) uint8, positives (L,) int64) from plain detection output
    over channels 0..M-1."""
    dev = times.device
    nW = (N + WINDOW - 1) // WINDOW
    labels = torch.empty((nW, M), dtype=torch.uint8, device=dev)
    positives = torch.empty(max(times.numel(), 1), dtype=torch.int64, device=dev)
    _capi.call("msamd_tcn_labels", _ptr(times), _ptr(offsets), int(M), int(N), _ptr(labels),
               _ptr(positives), _stream(dev))
    return labels, positives[:times.numel()]


def loss_and_grad(X, labels, params, seg_channel, seg_window, seg_windows, pos_weight=1.0):
    """Mean weighted BCE of the given segments and its gradient (fp32)."""
    _check_ts(X)
    N, M = X.shape
    dev = X.device
    seg_c = torch.as_tensor(seg_channel, dtype=torch.int32).to(dev).contiguous()
    seg_w = torch.as_tensor(seg_window, dtype=torch.int64).to(dev).contiguous()
    B = int(seg_c.numel())
    ws = _workspace(_capi.lib().msamd_tcn_train_workspace_size(B, int(seg_windows)), dev)
    grad = torch.empty(num_params(), dtype=torch.float32, device=dev)
    loss = torch.empty(1, dtype=torch.float32, device=dev)
    _capi.call("msamd_tcn_grad", _ptr(X), M, N, _ptr(labels), _ptr(params), _ptr(seg_c),
               _ptr(seg_w), B, int(seg_windows), float(pos_weight), _ptr(grad), _ptr(loss),
               _ptr(ws), ws.numel(), _stream(dev))
    return float(loss.item()), grad


def train(X, labels, positives, params, steps=4000, batch=256, seg_windows=16, lr=2e-3,
          pos_weight=20.0, seed=0, adam_state=None, step0=0):
    """Runs `steps` Adam steps on params in place; returns (loss per step, adam_state)."""
    _check_ts(X)
    N, M = X.shape
    dev = X.device
    if adam_state is None:
        adam_state = torch.zeros(2 * num_params(), dtype=torch.float32, device=dev)
    ws = _workspace(_capi.lib().msamd_tcn_train_workspace_size(int(batch), int(seg_windows)), dev)
    trace = torch.empty(max(int(steps), 1), dtype=torch.float32, device=dev)
    _capi.call("msamd_tcn_train", _ptr(X), M, N, _ptr(labels), _ptr(positives),
               int(positives.numel()), _ptr(params), _ptr(adam_state), int(steps), int(step0),
               int(batch), int(seg_windows), float(lr), float(pos_weight),
               ctypes.c_uint64(seed), _ptr(trace), _ptr(ws), ws.numel(), _stream(dev))
    return trace[:steps], adam_state


def score(X, params):
    """Logits (nW, M) of every window (bf16 MFMA inference kernel)."""
    _check_ts(X)
    N, M = X.shape
    dev = X.device
    nW = (N + WINDOW - 1) // WINDOW
    ws = _workspace(_capi.lib().msamd_tcn_score_workspace_size(), dev)
    z = torch.empty((nW, M), dtype=torch.float32, device=dev)
    _capi.call("msamd_tcn_score", _ptr(X), M, N, _ptr(params), _ptr(z), _ptr(ws), ws.numel(),
               _stream(dev))
    return z


class LearnedDetection:
    """Workspace-holding learned detection (reuse across calls of one shape)."""

    def __init__(self, params):
        self.params = params
        self._ws = None

    def __call__(self, X, detect_threshold=3.0, detect_interval=10, detect_sign=0,
                 num_threads=8, chunk_size=None, t1=-1, t2=-1, score_lo=0.2, score_hi=0.8,
                 padding=DETECT_PADDING):
        """Returns (offsets, times, class_counts) like frontend.detect_events on
        all channels, plus the windows per class {skipped, uncertain, present}."""
        _check_ts(X)
        N, M = X.shape
        dev = X.device
        t_start = 0 if t1 < 0 else int(t1)
        t_end = N - 1 if t2 < 0 else int(t2)
        n_range = t_end - t_start + 1
        cs = int(chunk_size) if chunk_size else detect_chunk_size(M, n_range, num_threads)
        L = _capi.lib()
        nb = L.msamd_tcn_detect_workspace_size(N, n_range, cs, M, float(detect_interval))
        if self._ws is None or self._ws.numel() < nb:
            self._ws = _workspace(nb, dev)
        ws = self._ws
        total = ctypes.c_int64(0)
        counts = (ctypes.c_int64 * 3)()
        st = _stream(dev)
        _capi.call("msamd_tcn_detect", _ptr(X), M, N, int(t1), int(t2), cs, int(padding),
                   _ptr(self.params), float(detect_threshold), float(detect_interval),
                   int(detect_sign), float(score_lo), float(score_hi), _ptr(ws), ws.numel(),
                   ctypes.byref(total), counts, st)
        times = torch.empty(max(total.value, 1), dtype=torch.int64, device=dev)
        offsets = torch.empty(M + 1, dtype=torch.int64, device=dev)
        _capi.call("msamd_detect_fetch", _ptr(ws), M, _ptr(times), _ptr(offsets), st)
        return offsets, times[:total.value], list(counts)


def train_detector(X, detect_threshold=3.0, detect_interval=10, detect_sign=-1, num_threads=8,
                   steps=4000, batch=256, seg_windows=16, lr=2e-3, pos_weight=20.0, seed=0):
    """Labels a training recording by plain detection and trains a fresh
    network on it.  Returns (params, loss trace)."""
    offsets, times = detect_events(X, detect_threshold, detect_interval, detect_sign,
                                   num_threads=num_threads)
    N, M = X.shape
    labels, positives = labels_from_events(offsets, times, M, N)
    params = init_params(seed, X.device)
    trace, _ = train(X, labels, positives, params, steps, batch, seg_windows, lr, pos_weight,
                     seed)
    return params, trace


def match_events(off_a, t_a, off_b, t_b, tol=1):
    """Per-channel matching of event times within +-tol samples (greedy, in time
    order).  Returns (matched, len_a, len_b): recall of b in a = matched/len_b."""
    off_a = off_a.cpu().tolist()
    off_b = off_b.cpu().tolist()
    ta = t_a.cpu()
    tb = t_b.cpu()
    matched = 0
    for c in range(len(off_a) - 1):
        a = ta[off_a[c]:off_a[c + 1]]
        b = tb[off_b[c]:off_b[c + 1]]
        if a.numel() == 0 or b.numel() == 0:
            continue
        # for every b, nearest a by searchsorted; count unique pairs within tol
        idx = torch.searchsorted(a, b)
        best = torch.full((b.numel(),), math.inf, dtype=torch.float64)
        for d in (-1, 0):
            j = (idx + d).clamp(0, a.numel() - 1)
            best = torch.minimum(best, (a[j] - b).abs().double())
        matched += int((best <= tol).sum())
    return matched, int(t_a.numel()), int(t_b.numel())
