"""GPU parity of the learned detector: training gradients against the FP64
oracle, inference logits against the oracle forward, the gated detection
scan bit-exact against the oracle's gated rule, and what training buys
(recall and precision against plain detection on a fresh recording)."""
import math

import numpy as np
import pytest
import torch

from oracle import tcn as ot
from tests.conftest import say

pytestmark = pytest.mark.gpu


def _bf16_weights(theta32):
    """Parameter vector as the kernels use it: weights rounded to bf16, biases fp32."""
    th = np.asarray(theta32, np.float32).astype(np.float64)
    P = ot.unpack(th)
    for n in ("W1", "W2", "W3"):
        P[n] = ot.bf16_round(P[n].astype(np.float32))
    return ot.pack(P)


def _spiky(N, M, seed, n_spikes=None, amp=(3, 9)):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(N, M))
    t = np.arange(50)
    for _ in range(n_spikes or N // 400):
        st = rng.integers(0, N - 50)
        c = rng.integers(0, M)
        w = -rng.uniform(*amp) * np.exp(-0.5 * ((t - 20) / 2) ** 2)
        for dc, g in ((0, 1.0), (-1, 0.5), (1, 0.5)):
            if 0 <= c + dc < M:
                X[st:st + 50, c + dc] += g * w
    return X.astype(np.float32)


def test_gradient_matches_oracle(gpu):
    from mountainsort_amd import learned
    X = _spiky(6000, 9, 0)
    N, M = X.shape
    nW = (N + 31) // 32
    rng = np.random.default_rng(1)
    labels = (rng.random((nW, M)) < 0.2).astype(np.uint8)
    theta = ot.init_params(3).astype(np.float32)
    theta[-1] = 0.0
    segs = [(0, 0), (4, 30), (8, 80), (2, nW - 5), (5, 100), (6, -3)]
    S = 5
    say(f"tcn grad: {len(segs)} segments of {S} windows, M={M} N={N}")
    loss_g, grad_g = learned.loss_and_grad(
        torch.from_numpy(X).to(gpu), torch.from_numpy(labels).to(gpu),
        torch.from_numpy(theta).to(gpu), [c for c, _ in segs], [w for _, w in segs], S, 3.0)
    loss_o, grad_o = ot.batch_loss_grad(_bf16_weights(theta), X, labels, segs, S, 3.0)
    grad_g = grad_g.cpu().numpy().astype(np.float64)
    e_loss = abs(loss_g - loss_o) / loss_o
    e_grad = np.max(np.abs(grad_g - grad_o)) / np.max(np.abs(grad_o))
    # per block of parameters too (the small layer-3 block must not hide in the max)
    blocks = {}
    o = 0
    for name, shape in ot.SIZES:
        n = int(np.prod(shape))
        ref = grad_o[o:o + n]
        blocks[name] = np.max(np.abs(grad_g[o:o + n] - ref)) / max(np.max(np.abs(ref)), 1e-30)
        o += n
    say(f"  loss rel err {e_loss:.2e}, grad max rel err {e_grad:.2e}, per block "
        + ", ".join(f"{k} {v:.1e}" for k, v in blocks.items()))
    assert e_loss < 1e-5
    assert e_grad < 1e-4
    assert max(blocks.values()) < 1e-3


def test_inference_matches_oracle(gpu):
    from mountainsort_amd import learned
    X = _spiky(20_000, 11, 2)
    theta = ot.init_params(4).astype(np.float32)
    theta[-1] = 0.0
    say("tcn score: M=11 N=20000")
    z_g = learned.score(torch.from_numpy(X).to(gpu), torch.from_numpy(theta).to(gpu))
    z_g = z_g.cpu().numpy().astype(np.float64)
    # the inference kernel rounds the input samples and the layer-1 activations
    # to bf16; the oracle applies the same input rounding
    Xb = ot.bf16_round(X).astype(np.float32)
    z_o = ot.logits_all(_bf16_weights(theta), Xb)
    err = np.abs(z_g - z_o)
    scale = np.max(np.abs(z_o))
    say(f"  logits: max |diff| {err.max():.3e}, scale {scale:.2f}, rel {err.max() / scale:.2e}")
    assert z_g.shape == z_o.shape
    assert err.max() < 1e-2 * scale


@pytest.mark.parametrize("sign,D,t1,t2", [(-1, 10, -1, -1), (0, 3, 5000, 150_000)])
def test_gated_detection_matches_oracle(gpu, sign, D, t1, t2):
    from mountainsort_amd import learned
    X = _spiky(200_000, 12, 5)
    theta = ot.init_params(6).astype(np.float32)
    Xd = torch.from_numpy(X).to(gpu)
    pd = torch.from_numpy(theta).to(gpu)
    z = learned.score(Xd, pd).cpu().numpy()
    # band chosen so that all three classes occur
    q_lo, q_hi = np.quantile(z, [0.4, 0.8])
    lo, hi = 1 / (1 + math.exp(-q_lo)), 1 / (1 + math.exp(-q_hi))
    zl = np.float32(math.log(lo / (1 - lo)))
    zh = np.float32(math.log(hi / (1 - hi)))
    cls = np.where(z < zl, 0, np.where(z > zh, 2, 1)).astype(np.uint8)
    say(f"gated detection sign={sign} D={D} t1={t1} t2={t2}: classes "
        f"{[int((cls == k).sum()) for k in range(3)]}")
    det = learned.LearnedDetection(pd)
    off, times, counts = det(Xd, 3.0, D, sign, chunk_size=20_000, t1=t1, t2=t2, score_lo=lo,
                             score_hi=hi)
    cnt_o, t_o = ot.gated_detect(X, cls, 3.0, D, sign, chunk_size=20_000, t1=t1, t2=t2)
    off = off.cpu().numpy()
    assert np.array_equal(np.diff(off), cnt_o)
    assert np.array_equal(times.cpu().numpy(), t_o)
    w0 = (0 if t1 < 0 else t1) // 32
    w1 = (X.shape[0] - 1 if t2 < 0 else t2) // 32 + 1
    assert counts == [int((cls[w0:w1] == k).sum()) for k in range(3)]


def test_training_gain_over_untrained(gpu):
    """Train on one recording, evaluate on another: recall of plain detection's
    events (within +-1 sample) and precision, against an untrained network."""
    from mountainsort_amd import frontend as fe
    from mountainsort_amd import learned
    from mountainsort_amd import synth

    def whitened(seed, N):
        cfg = synth.config1(N=N, seed=seed)
        cfg.M = 64
        cfg.num_units = 48
        raw = synth.generate(cfg, gpu)
        filt = fe.bandpass_filter(raw, 30000, 300, 6000, num_threads=8)
        return fe.whiten(filt)

    say("tcn training: 64 ch x 60 s, 4000 steps; evaluation on a fresh 64 ch x 20 s")
    Xtr = whitened(11, 1_800_000)
    params, trace = learned.train_detector(Xtr, 3.0, 10, -1, steps=4000, seed=1)
    trace = trace.cpu().numpy()
    say(f"  loss: first 50 mean {trace[:50].mean():.4f}, last 50 mean {trace[-50:].mean():.4f}")
    assert trace[-50:].mean() < 0.5 * trace[:50].mean()

    Xte = whitened(12, 600_000)
    off_p, t_p = fe.detect_events(Xte, 3.0, 10, -1, num_threads=8)
    results = {}
    for name, p in (("trained", params), ("untrained", learned.init_params(1, gpu))):
        off_l, t_l, counts = learned.LearnedDetection(p)(Xte, 3.0, 10, -1, num_threads=8)
        hit_p, _, n_p = learned.match_events(off_l, t_l, off_p, t_p)   # plain events found
        hit_l, _, n_l = learned.match_events(off_p, t_p, off_l, t_l)   # learned events confirmed
        recall = hit_p / max(n_p, 1)
        precision = hit_l / max(n_l, 1)
        results[name] = (recall, precision, counts)
        say(f"  {name}: recall {recall:.4f} precision {precision:.4f} "
            f"({n_l} vs {n_p} plain events), windows skipped/uncertain/present {counts}")
    rec, prec, counts = results["trained"]
    assert rec >= 0.99
    assert prec >= 0.99
    # the trained network must let detection skip most windows
    assert counts[0] > 0.8 * sum(counts)
    assert results["untrained"][0] < rec or results["untrained"][2][0] < counts[0]
