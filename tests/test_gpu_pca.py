"""GPU parity: batched two-stage PCA features against the CPU oracle (the
reference's float-form PCA) and the FP64 Gram restatement."""
import numpy as np
import pytest
import torch

from oracle import core as oc
from tests.conftest import say

pytestmark = pytest.mark.gpu


def _clips(rng, L, W, T=50):
    """Spike-like clips: a few templates with per-event amplitude and noise."""
    t = np.arange(T)
    K = 4
    temps = []
    for _ in range(K):
        w = -rng.uniform(3, 12) * np.exp(-0.5 * ((t - 24) / rng.uniform(1.5, 3)) ** 2) \
            + rng.uniform(0.5, 3) * np.exp(-0.5 * ((t - 32) / 5) ** 2)
        temps.append(np.outer(w, rng.uniform(0.1, 1, W)))
    if L == 0:
        return np.zeros((0, T, W), np.float32)
    lab = rng.integers(0, K, L)
    amp = rng.uniform(0.7, 1.3, L)
    X = np.stack([temps[k] for k in lab]) * amp[:, None, None] + rng.normal(size=(L, T, W))
    return X.astype(np.float32)  # (L, T, W)


def test_pca_features_match_oracle(gpu):
    from mountainsort_amd import frontend as fe
    rng = np.random.default_rng(3)
    shapes = [(300, 5), (0, 3), (12345, 12), (50, 1), (1500, 16)]
    say(f"pca neighborhoods (L, W) = {shapes}")
    clips = [_clips(rng, L, W) for L, W in shapes]
    flat = torch.from_numpy(np.concatenate([c.ravel() for c in clips])).to(gpu)
    eo = np.concatenate([[0], np.cumsum([L for L, _ in shapes])])
    no = np.concatenate([[0], np.cumsum([W for _, W in shapes])])
    co = np.concatenate([[0], np.cumsum([L * W * 50 for L, W in shapes])])
    red, ro, feat, fo = fe.pca_features(flat, co, eo, no, 50)
    red = red.cpu().numpy()
    feat = feat.cpu().numpy()
    worst1 = worst2 = 0.0
    for g, ((L, W), c) in enumerate(zip(shapes, clips)):
        if L == 0:
            continue
        r_o = oc.dimension_reduce_clips(c, 10, 10000)              # (L, 10, W)
        f_o = oc.stage2_features(r_o, 10, 10000)                   # (L, 10)
        r_g = red[ro[g]:ro[g + 1]].reshape(L, 10, W)
        f_g = feat[fo[g]:fo[g + 1]].reshape(L, 10)
        # components are defined up to the power method's own sign choice;
        # both start from sin(1..n), so signs agree and values compare directly
        e1 = np.max(np.abs(r_g - r_o)) / np.max(np.abs(r_o))
        e2 = np.max(np.abs(f_g - f_o)) / np.max(np.abs(f_o))
        # the first components are well separated: tighter check on them
        e1a = np.max(np.abs(r_g[:, :3] - r_o[:, :3])) / np.max(np.abs(r_o[:, :3]))
        say(f"  g={g} L={L} W={W}: stage1 max rel {e1:.2e} (first 3: {e1a:.2e}), stage2 {e2:.2e}")
        worst1, worst2 = max(worst1, e1a), max(worst2, e2)
        assert e1a < 1e-3
    assert worst1 < 1e-3
