"""FP64 oracle of the learned spike-presence detector (test infrastructure only).

The detector has no counterpart in the reference; its definition is this
project's own, and this module is the precision reference for the GPU
kernels: forward pass, weighted binary cross-entropy, analytic gradients (pinned
by central finite differences in tests/test_oracle_tcn.py) and the gated
detection rule that turns window scores back into event times with the
reference's detection rule (detect_events.cpp:69-135 as restated in
oracle.core) as the fallback.

Network (all index conventions exact):
  input   x[ci][t] = X[t, c - 3 + ci], ci = 0..6 (zero outside the recording
          and outside [0, M)): the central channel c and its three lower and
          three higher channel indices
  layer 1 h1[o][j] = relu(b1[o] + sum_{ci,k} W1[o][ci][k] x[ci][8j + 4 + k])
  layer 2 h2[o][p] = relu(b2[o] + sum_{i,k}  W2[o][i][k]  h1[i][4p - 3 + k])
  layer 3 z[w]     = b3 + sum_{i,k} W3[i][k] h2[i][w - 4 + k]
  o, i = 0..31, k = 0..8.  z[w] scores window w = samples [32w, 32w + 32) of
  channel c (receptive field centred on sample 32w + 16).
Parameters are one float vector: W1 (32*7*9) | b1 (32) | W2 (32*32*9) |
b2 (32) | W3 (32*9) | b3 (1).
"""
import numpy as np

from oracle import core as oc

H, CI, K = 32, 7, 9
WIN = 32
SIZES = [("W1", (H, CI, K)), ("b1", (H,)), ("W2", (H, H, K)), ("b2", (H,)), ("W3", (H, K)),
         ("b3", (1,))]
NUM_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(theta):
    theta = np.asarray(theta, np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = theta[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[n], np.float64).ravel() for n, _ in SIZES])


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float64."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def segment_input(X, c, w0, S):
    """x[ci][t] for the samples a segment of S windows starting at window w0 needs.
    Returns (x (7, nt) float64, t0) with x[:, t - t0]."""
    X = np.asarray(X)
    N, M = X.shape
    jlo = 4 * (w0 - 4) - 3
    jhi = 4 * (w0 + S + 3) + 5
    t0, t1 = 8 * jlo + 4, 8 * jhi + 12
    x = np.zeros((CI, t1 - t0 + 1))
    a, b = max(t0, 0), min(t1 + 1, N)
    for ci in range(CI):
        m = c - 3 + ci
        if 0 <= m < M and a < b:
            x[ci, a - t0:b - t0] = X[a:b, m]
    return x, t0


def forward_segment(P, x, t0, w0, S):
    """Returns (z (S,), cache) for windows w0..w0+S-1."""
    plo, phi = w0 - 4, w0 + S + 3
    jlo, jhi = 4 * plo - 3, 4 * phi + 5
    J = np.arange(jlo, jhi + 1)
    # layer 1
    idx1 = (8 * J[:, None] + 4 + np.arange(K)[None, :]) - t0          # (nj, K)
    patches1 = x[:, idx1]                                               # (CI, nj, K)
    pre1 = np.einsum("oik,ijk->oj", P["W1"], patches1) + P["b1"][:, None]
    h1 = np.maximum(pre1, 0.0)
    # layer 2
    Pp = np.arange(plo, phi + 1)
    idx2 = (4 * Pp[:, None] - 3 + np.arange(K)[None, :]) - jlo          # (np, K)
    patches2 = h1[:, idx2]                                              # (H, np, K)
    pre2 = np.einsum("oik,ipk->op", P["W2"], patches2) + P["b2"][:, None]
    h2 = np.maximum(pre2, 0.0)
    # layer 3
    Wn = np.arange(w0, w0 + S)
    idx3 = (Wn[:, None] - 4 + np.arange(K)[None, :]) - plo              # (S, K)
    z = np.einsum("ik,iwk->w", P["W3"], h2[:, idx3]) + P["b3"][0]
    cache = dict(patches1=patches1, pre1=pre1, h1=h1, idx2=idx2, patches2=patches2,
                 pre2=pre2, h2=h2, idx3=idx3, jlo=jlo, plo=plo)
    return z, cache


def bce_grad(z, y, pos_weight):
    """Weighted BCE with logits: loss per window and d loss / d z."""
    z = np.asarray(z, np.float64)
    y = np.asarray(y, np.float64)
    sp_pos = np.logaddexp(0.0, -z)   # -log sigmoid(z)
    sp_neg = np.logaddexp(0.0, z)    # -log(1 - sigmoid(z))
    loss = pos_weight * y * sp_pos + (1.0 - y) * sp_neg
    s = 1.0 / (1.0 + np.exp(-z))
    dz = pos_weight * y * (s - 1.0) + (1.0 - y) * s
    return loss, dz


def backward_segment(P, cache, dz):
    """Gradients of sum(dz * z) with respect to every parameter (dict)."""
    h2, idx3 = cache["h2"], cache["idx3"]
    g = {}
    g["W3"] = np.einsum("w,iwk->ik", dz, h2[:, idx3])
    g["b3"] = np.array([dz.sum()])
    dh2 = np.zeros_like(h2)
    np.add.at(dh2, (slice(None), idx3), P["W3"][:, None, :] * dz[None, :, None])
    dpre2 = dh2 * (cache["pre2"] > 0)
    g["W2"] = np.einsum("op,ipk->oik", dpre2, cache["patches2"])
    g["b2"] = dpre2.sum(axis=1)
    h1 = cache["h1"]
    dh1 = np.zeros_like(h1)
    contrib = np.einsum("oik,op->ipk", P["W2"], dpre2)                  # (H, np, K)
    np.add.at(dh1, (slice(None), cache["idx2"]), contrib)
    dpre1 = dh1 * (cache["pre1"] > 0)
    g["W1"] = np.einsum("oj,ijk->oik", dpre1, cache["patches1"])
    g["b1"] = dpre1.sum(axis=1)
    return g


def batch_loss_grad(theta, X, labels, segs, S, pos_weight):
    """Mean weighted BCE over every window of every segment (c, w0) and its
    gradient.  labels: (nW, M) 0/1 array (window w of channel c at [w, c];
    windows outside [0, nW) count as 0)."""
    P = unpack(theta)
    nW = labels.shape[0]
    total = 0.0
    grad = {n: np.zeros(s) for n, s in SIZES}
    for c, w0 in segs:
        x, t0 = segment_input(X, c, w0, S)
        z, cache = forward_segment(P, x, t0, w0, S)
        w = np.arange(w0, w0 + S)
        y = np.where((w >= 0) & (w < nW), labels[np.clip(w, 0, nW - 1), c], 0)
        loss, dz = bce_grad(z, y, pos_weight)
        total += loss.sum()
        g = backward_segment(P, cache, dz)
        for n in grad:
            grad[n] += g[n]
    n = len(segs) * S
    return total / n, pack(grad) / n


def logits_all(theta, X):
    """z for every window of every channel: (nW, M) with nW = ceil(N / 32)."""
    X = np.asarray(X)
    N, M = X.shape
    nW = (N + WIN - 1) // WIN
    P = unpack(theta)
    out = np.empty((nW, M))
    for c in range(M):
        x, t0 = segment_input(X, c, 0, nW)
        out[:, c], _ = forward_segment(P, x, t0, 0, nW)
    return out


def window_labels(counts, times, M, N):
    """(nW, M) uint8: 1 where the plain detection has an event in window w of channel c."""
    nW = (N + WIN - 1) // WIN
    lab = np.zeros((nW, M), np.uint8)
    ch = np.repeat(np.arange(M), counts)
    lab[np.asarray(times) // WIN, ch] = 1
    return lab


def gated_detect(X, classes, detect_threshold, detect_interval, detect_sign, num_threads=8,
                 chunk_size=None, t1=-1, t2=-1, padding=1000):
    """Event times of the gated rule on every channel, per channel ascending.

    classes (nW, M): 0 = no spike (the window is skipped); 1 = uncertain and
    2 = present (the window was proposed by the network): the plain rule of
    detect_events decides every sample.  Chunk plan, statistics and v are
    those of ms3alg detection (p_mountainsort3.cpp:170-216,507-533;
    detect_events.cpp).
    Returns (counts (M,), times concatenated by channel)."""
    X = np.asarray(X, np.float32)
    N, M = X.shape
    t_start = 0 if t1 < 0 else int(t1)
    t_end = N - 1 if t2 < 0 else int(t2)
    n_range = t_end - t_start + 1
    cs = int(chunk_size) if chunk_size else oc.detect_chunk_size(M, n_range, num_threads)
    D = int(np.floor(detect_interval)) if detect_interval >= 0 else -1
    counts = np.zeros(M, np.int64)
    out = []
    nchunks = (n_range + cs - 1) // cs
    for c in range(M):
        tc = []
        for ck in range(nchunks):
            core0 = t_start + ck * cs
            size = min(cs, n_range - ck * cs)
            wa, wb = core0 - padding, core0 + size + padding
            seg = np.zeros(wb - wa)
            a, b = max(wa, 0), min(wb, N)
            seg[a - wa:b - wa] = X[a:b, c]
            mean, sd = oc.mean_stdev(seg)
            thr = detect_threshold * sd
            v = seg - mean
            if detect_sign < 0:
                v = -v
            elif detect_sign == 0:
                v = np.abs(v)
            n = len(v)
            keep = v > thr
            for d in range(1, D + 1):
                left = np.full(n, -np.inf)
                left[d:] = v[:-d]
                right = np.full(n, -np.inf)
                right[:-d] = v[d:]
                keep &= (v > left) & (v >= right)
            t = np.arange(core0, core0 + size)
            sel = keep[t - wa] & (classes[t // WIN, c] != 0)
            tc.extend(t[sel].tolist())
        counts[c] = len(tc)
        out.extend(tc)
    return counts, np.asarray(out, np.int64)


def init_params(seed=0):
    """He-normal weights, zero biases, b3 at the log-odds of a 5% positive rate."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(size=(H, CI, K)) * np.sqrt(2.0 / (CI * K)), "b1": np.zeros(H),
         "W2": rng.normal(size=(H, H, K)) * np.sqrt(2.0 / (H * K)), "b2": np.zeros(H),
         "W3": rng.normal(size=(H, K)) * np.sqrt(1.0 / (H * K)), "b3": np.array([-3.0])}
    return pack(d)
