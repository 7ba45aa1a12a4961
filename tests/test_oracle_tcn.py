"""The FP64 oracle of the learned detector: analytic gradients against central
finite differences, window bookkeeping, and the gated detection rule against
the plain detection oracle.  CPU only."""
import numpy as np
import pytest

from oracle import core as oc
from oracle import tcn as ot


def _data(N=3000, M=9, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(N, M)).astype(np.float32)
    for st in rng.integers(0, N - 50, 12):
        X[st + 18:st + 24, rng.integers(0, M)] -= 6.0
    return X


def test_num_params():
    assert ot.NUM_PARAMS == 32 * 7 * 9 + 32 + 32 * 32 * 9 + 32 + 32 * 9 + 1 == 11585


def test_gradient_matches_finite_differences():
    X = _data()
    M = X.shape[1]
    nW = (X.shape[0] + 31) // 32
    rng = np.random.default_rng(1)
    labels = (rng.random((nW, M)) < 0.2).astype(np.uint8)
    theta = ot.init_params(3)
    theta[-1] = 0.0   # b3: balanced logits so both loss branches contribute
    segs = [(0, 0), (4, 30), (8, 80), (2, nW - 5)]   # includes both edges of the recording
    S = 5
    loss, g = ot.batch_loss_grad(theta, X, labels, segs, S, pos_weight=3.0)
    idx = np.concatenate([rng.choice(ot.NUM_PARAMS - 1, 60, replace=False), [ot.NUM_PARAMS - 1]])
    h = 1e-6
    worst = 0.0
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (ot.batch_loss_grad(tp, X, labels, segs, S, 3.0)[0] -
              ot.batch_loss_grad(tm, X, labels, segs, S, 3.0)[0]) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / (abs(fd) + 1e-7))
    assert worst < 1e-4, worst


def test_zero_weights_give_bias_logits():
    X = _data(N=640, M=3)
    theta = np.zeros(ot.NUM_PARAMS)
    theta[-1] = 0.7
    z = ot.logits_all(theta, X)
    assert z.shape == (20, 3) and np.all(z == 0.7)


def test_receptive_field_is_local():
    """A change at sample t moves only windows whose receptive field
    [32w - 148, 32w + 180] contains t, and only channels c with |c - m| <= 3."""
    X = _data(N=4096, M=12)
    theta = ot.init_params(5)
    z0 = ot.logits_all(theta, X)
    X2 = X.copy()
    X2[2000, 6] += 50.0
    dz = np.abs(ot.logits_all(theta, X2) - z0) > 0
    w, c = np.nonzero(dz)
    assert len(w) > 0
    assert np.all(np.abs(c - 6) <= 3)
    assert np.all((32 * w - 148 <= 2000) & (2000 <= 32 * w + 180))


def test_labels_mark_event_windows():
    lab = ot.window_labels([2, 1], [5, 40, 31], M=2, N=100)
    assert lab.shape == (4, 2)
    assert lab[0, 0] == 1 and lab[1, 0] == 1 and lab[0, 1] == 1 and lab.sum() == 3


def _whitened(M=6, N=60_000, seed=2):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(N, M))
    t = np.arange(50)
    for _ in range(120):
        st = rng.integers(0, N - 50)
        X[st:st + 50, rng.integers(0, M)] += -rng.uniform(3, 9) * np.exp(-0.5 * ((t - 20) / 2) ** 2)
    return X.astype(np.float32)


@pytest.mark.parametrize("sign,D,cs", [(-1, 10, 20_000), (0, 3, 20_000), (1, 0, 25_000)])
def test_gated_rule_all_uncertain_equals_plain_detection(sign, D, cs):
    X = _whitened()
    nW = (X.shape[0] + 31) // 32
    cls = np.ones((nW, X.shape[1]), np.uint8)
    cnt, t = ot.gated_detect(X, cls, 3.0, D, sign, chunk_size=cs)
    cnt_p, t_p = oc.ms3alg_detect(X, 3.0, D, sign, chunk_size=cs)
    assert np.array_equal(cnt, cnt_p) and np.array_equal(t, t_p)


def test_gated_rule_skip_and_proposed_classes():
    X = _whitened()
    N, M = X.shape
    nW = (N + 31) // 32
    cnt, _ = ot.gated_detect(X, np.zeros((nW, M), np.uint8), 3.0, 10, -1, chunk_size=20_000)
    assert cnt.sum() == 0
    cnt_p, t_p = oc.ms3alg_detect(X, 3.0, 10, -1, chunk_size=20_000)
    cnt2, t2 = ot.gated_detect(X, np.full((nW, M), 2, np.uint8), 3.0, 10, -1, chunk_size=20_000)
    assert np.array_equal(cnt2, cnt_p) and np.array_equal(t2, t_p)
    # a random mixture: exactly the plain events that fall in proposed windows
    rng = np.random.default_rng(7)
    cls = rng.integers(0, 3, (nW, M)).astype(np.uint8)
    cnt3, t3 = ot.gated_detect(X, cls, 3.0, 10, -1, chunk_size=20_000)
    off = np.concatenate([[0], np.cumsum(cnt_p)])
    ref = []
    for c in range(M):
        tc = t_p[off[c]:off[c + 1]]
        ref.append(tc[cls[tc // 32, c] != 0])
    assert np.array_equal(t3, np.concatenate(ref))
    assert np.array_equal(cnt3, [len(r) for r in ref])
