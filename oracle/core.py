"""Numpy restatements and ctypes bindings of the C restatement (oracle/msoracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's
CPU-baseline leg, never by the product path.

Array conventions follow the reference's .mda column-major order: a timeseries
M x N is held here as a numpy array of shape (N, M) in C order, so the memory
is ``x[m + M*t]`` exactly as in ``Mda32`` (packages/mountainsortalg/mda/mda32.h).
Clips W x T x L are held as shape (L, T, W).
"""
import ctypes
import os

import numpy as np
import scipy.fft as _sfft

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libmsoracle.so")
        if not os.path.exists(path):
            from oracle import build as _b
            _b.build()
        L = ctypes.CDLL(path)
        i64, f64, i32, vp = ctypes.c_int64, ctypes.c_double, ctypes.c_int32, ctypes.c_void_p
        L.orc_define_kernel.argtypes = [i64, vp, f64, f64, f64, f64]
        L.orc_bandpass_chunk_size.argtypes = [i64, i64]
        L.orc_bandpass_chunk_size.restype = i64
        L.orc_detect_chunk_size.argtypes = [i64, i64, i64]
        L.orc_detect_chunk_size.restype = i64
        L.orc_mean_stdev.argtypes = [vp, i64, vp, vp]
        L.orc_detect_events.argtypes = [vp, i64, f64, f64, ctypes.c_int, vp]
        L.orc_detect_events.restype = i64
        L.orc_ms3alg_detect.argtypes = [vp, i64, i64, i64, i64, i64, i64, vp, i32, f64, f64,
                                        ctypes.c_int, vp, vp, i64]
        L.orc_ms3alg_detect.restype = i64
        L.orc_whiten_covariance.argtypes = [vp, i64, i64, vp]
        L.orc_pca_from_XXt.argtypes = [vp, i64, i64, vp, vp]
        L.orc_whitening_matrix.argtypes = [vp, i64, vp]
        L.orc_whiten_apply.argtypes = [vp, i64, i64, vp, vp, f64, vp]
        L.orc_whiten_clips.argtypes = [vp, i64, i64, vp, vp, f64, vp]
        L.orc_mask_out_artifacts.argtypes = [vp, i64, i64, f64, i64, vp, vp, vp]
        L.orc_mask_out_artifacts.restype = ctypes.c_int
        L.orc_pca.argtypes = [vp, i64, i64, i64, vp, vp, vp]
        L.orc_pca_subsample_indices.argtypes = [i64, i64, vp]
        L.orc_pca_subsample_indices.restype = i64
        L.orc_pca_subsampled.argtypes = [vp, i64, i64, i64, i64, vp, vp, vp]
        L.orc_dimension_reduce_clips.argtypes = [vp, i64, i64, i64, i64, i64, vp]
        L.orc_compute_templates.argtypes = [vp, i64, i64, vp, vp, i64, i32, i32, vp, vp]
        L.orc_fit_stage.argtypes = [vp, i64, i64, vp, vp, i64, vp, i32, i32, f64, vp]
        L.orc_fit_stage.restype = i64
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------
# a1-a3: bandpass filter (packages/ms3/p_bandpass_filter.cpp)
# --------------------------------------------------------------------------
def define_kernel(n_fft, samplerate, freq_min, freq_max, freq_wid):
    """Gain H[i] of define_kernel, p_bandpass_filter.cpp:421-447 (C restatement)."""
    out = np.empty(int(n_fft), np.float64)
    lib().orc_define_kernel(int(n_fft), _p(out), float(samplerate), float(freq_min),
                            float(freq_max), float(freq_wid))
    return out


def bandpass_chunk_size(M, num_threads):
    """Core chunk length of the filter windows, p_bandpass_filter.cpp:147-155."""
    return int(lib().orc_bandpass_chunk_size(int(M), int(num_threads)))


BANDPASS_OVERLAP = 60000  # p_bandpass_filter.cpp:150


def bandpass_filter(X, samplerate, freq_min, freq_max, freq_wid=1000.0, num_threads=8,
                    chunk_size=None, subsample_factor=1, quantization_unit=0.0,
                    overlap=BANDPASS_OVERLAP):
    """FP64 restatement of ms3.bandpass_filter (p_bandpass_filter.cpp:115-229).

    X: (N, M) array (any real dtype; read as float32 like DiskReadMda32).
    Each window [t-60000, t+chunk+60000) is zero-filled outside [0, N), sent
    through a length n_fft = chunk+120000 DFT, scaled by H/n_fft, inverse
    DFT'd, and the core [60000, 60000+chunk) kept (:178-199).  Arithmetic is
    FP64 throughout (the reference uses FFTW single precision; this is the
    FP64 oracle of that computation).  Returns float64 (N/k, M) or the int16
    file values when quantization_unit is set.  `overlap` is 60000 in the
    reference; smaller values exist for brute-force checks on tiny inputs.
    """
    X = np.asarray(X)
    N, M = X.shape
    if freq_max == 0:
        return X.copy()  # :117-119 copies the file
    k = int(subsample_factor) if subsample_factor else 1
    ov = int(overlap)
    cs = int(chunk_size) if chunk_size else bandpass_chunk_size(M, num_threads)
    n_fft = cs + 2 * ov
    H = define_kernel(n_fft, samplerate, freq_min, freq_max, freq_wid)
    N2 = N // k
    out = np.zeros((N2, M), np.float64)
    Xf = X.astype(np.float32).astype(np.float64)
    for t in range(0, N, cs):
        win = np.zeros((n_fft, M), np.float64)
        a, b = t - ov, t - ov + n_fft
        sa, sb = max(a, 0), min(b, N)
        if sb > sa:
            win[sa - a:sb - a] = Xf[sa:sb]
        # FP64 DFT pair (pocketfft through scipy, split over the host's cores)
        Y = _sfft.ifft(_sfft.fft(win, axis=0, workers=-1) * H[:, None], axis=0,
                       workers=-1).real
        core = Y[ov:ov + cs]
        if k > 1:  # subsample :204-206, :480-491
            core = core[:(cs // k) * k:k]
        i0 = t // k
        n = min(core.shape[0], N2 - i0)  # writeChunk clamps at the end, diskwritemda.cpp:243-245
        if n > 0:
            out[i0:i0 + n] = core[:n]
    if quantization_unit:
        # :207-209 then float->int16 conversion (truncation) on write, mdaio.cpp:247-258
        scaled = (out.astype(np.float32).astype(np.float64) * (1.0 / quantization_unit)).astype(np.float32)
        return np.trunc(scaled).astype(np.int16)
    return out


# --------------------------------------------------------------------------
# a4-a6: whitening (packages/ms3/p_whiten.cpp, packages/mountainsortalg/pca.cpp)
# --------------------------------------------------------------------------
def whiten_covariance(X):
    """p_whiten pass 1 (p_whiten.cpp:64-118): float products, double sums, /(N-1)."""
    X = np.ascontiguousarray(X, np.float32)
    N, M = X.shape
    C = np.empty((M, M), np.float64)
    lib().orc_whiten_covariance(_p(X), M, N, _p(C))
    return C  # symmetric; storage order irrelevant


def covariance_fp64(X):
    """Exact-product FP64 covariance X X^T / (N-1) (closed form for a4)."""
    Xd = np.asarray(X, np.float64)
    N = Xd.shape[0]
    C = Xd.T @ Xd
    return C / (N - 1) if N > 1 else C


def pca_from_XXt(C, K):
    """pca_from_XXt (pca.cpp:171-193), double form.  Returns (components M x K, sigma K)."""
    C = np.ascontiguousarray(np.asarray(C, np.float64).T)  # column-major storage
    M = C.shape[0]
    comp = np.empty(M * K, np.float64)
    sig = np.empty(K, np.float64)
    lib().orc_pca_from_XXt(_p(C), M, K, _p(comp), _p(sig))
    return comp.reshape(K, M).T.copy(), sig


def whitening_matrix(C):
    """whitening_matrix_from_XXt (pca.cpp:634-650), naive-deflation power method."""
    Cc = np.ascontiguousarray(np.asarray(C, np.float64).T)
    M = Cc.shape[0]
    W = np.empty(M * M, np.float64)
    lib().orc_whitening_matrix(_p(Cc), M, _p(W))
    return W.reshape(M, M).T.copy()


def whiten_apply(X, W, quantization_unit=0.0):
    """p_whiten pass 2 (p_whiten.cpp:153-206): float accumulation, quantize 1e-4.

    Returns float32 (N, M); with quantization_unit > 0 returns (float32, int16).
    """
    X = np.ascontiguousarray(X, np.float32)
    N, M = X.shape
    Wc = np.ascontiguousarray(np.asarray(W, np.float64).T)
    Y = np.empty((N, M), np.float32)
    Yq = np.empty((N, M), np.int16) if quantization_unit > 0 else None
    lib().orc_whiten_apply(_p(X), M, N, _p(Wc), _p(Y), float(quantization_unit),
                           _p(Yq) if Yq is not None else None)
    return (Y, Yq) if Yq is not None else Y


def whiten_clips(clips, W, quantization_unit=0.0):
    """ms3.whiten_clips (p_whiten.cpp:295-345): every (clip, time sample) column
    of the (L, T, M) clips times W with the float accumulation of pass 2 and no
    1e-4 quantize.  Returns float32 (L, T, M); with quantization_unit > 0
    returns (float32, int16 file values)."""
    X = np.ascontiguousarray(clips, np.float32)
    L, T, M = X.shape
    Wc = np.ascontiguousarray(np.asarray(W, np.float64).T)
    Y = np.empty_like(X)
    Yq = np.empty(X.shape, np.int16) if quantization_unit > 0 else None
    lib().orc_whiten_clips(_p(X), M, L * T, _p(Wc), _p(Y), float(quantization_unit),
                           _p(Yq) if Yq is not None else None)
    return (Y, Yq) if Yq is not None else Y


def quantize_1e4(Y):
    """quantize(x, 1e-4) of p_whiten.cpp:31-34 applied to float32 values."""
    Yd = np.asarray(Y, np.float32).astype(np.float64)
    return (np.floor(Yd / 0.0001 + 0.5) * 0.0001).astype(np.float32)


def whiten_apply_fp64(X, W):
    """FP64 apply W x (no quantization): the precision reference for a6."""
    return np.asarray(X, np.float64) @ np.asarray(W, np.float64).T


# --------------------------------------------------------------------------
# a7: mask_out_artifacts (packages/ms3/p_mask_out_artifacts.cpp:24-104)
# --------------------------------------------------------------------------
def mask_out_artifacts(X, threshold=6.0, interval_size=2000):
    """Returns (Y float32 (N, M), norms (nint, M) float64, use_it int32 (nint+1,))."""
    X = np.ascontiguousarray(X, np.float32)
    N, M = X.shape
    nint = N // int(interval_size)
    Y = np.empty((N, M), np.float32)
    norms = np.empty(max(M * nint, 1), np.float64)
    use_it = np.empty(nint + 1, np.int32)
    rc = lib().orc_mask_out_artifacts(_p(X), M, N, float(threshold), int(interval_size), _p(Y),
                                      _p(norms), _p(use_it))
    if rc != 0:
        raise ValueError("threshold or interval_size is zero")
    return Y, norms[:M * nint].reshape(nint, M), use_it


# --------------------------------------------------------------------------
# a8-a10: detection (packages/mountainsortalg)
# --------------------------------------------------------------------------
DETECT_PADDING = 1000  # p_mountainsort3.cpp:515


def detect_chunk_size(M, N, num_threads):
    """get_time_chunk_infos chunk length (p_mountainsort3.cpp:170-172,191,507-533)."""
    return int(lib().orc_detect_chunk_size(int(M), int(N), int(num_threads)))


def detect_events_sm(x, detect_threshold, detect_interval, sign):
    """Literal state machine of detect_events (detect_events.cpp:69-135) on a double vector."""
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty(max(len(x), 1), np.int64)
    n = lib().orc_detect_events(_p(x), len(x), float(detect_threshold), float(detect_interval),
                                int(sign), _p(out))
    return out[:n].copy()


def mean_stdev(x):
    x = np.ascontiguousarray(x, np.float64)
    m = ctypes.c_double()
    s = ctypes.c_double()
    lib().orc_mean_stdev(_p(x), len(x), ctypes.byref(m), ctypes.byref(s))
    return m.value, s.value


def detect_events_closed_form(x, detect_threshold, detect_interval, sign):
    """Closed form of the same rule: keep n iff v[n] > thr, v[n] > max v[n-D..n-1]
    and v[n] >= max v[n+1..n+D] (windows truncated at the array ends), with
    v = x - mean (negated for sign < 0, |.| for sign 0) and thr = k * stdev.
    Independent of the state-machine formulation; used to pin it."""
    x = np.asarray(x, np.float64)
    mean, sd = mean_stdev(x)
    thr = detect_threshold * sd
    v = x - mean
    if sign < 0:
        v = -v
    elif sign == 0:
        v = np.abs(v)
    D = int(np.floor(detect_interval))
    n = len(v)
    keep = v > thr
    for d in range(1, D + 1):
        left = np.full(n, -np.inf)
        left[d:] = v[:-d]
        right = np.full(n, -np.inf)
        right[:-d] = v[d:]
        keep &= (v > left) & (v >= right)
    return np.nonzero(keep)[0].astype(np.int64)


def ms3alg_detect(X, detect_threshold, detect_interval, detect_sign, num_threads=8,
                  chunk_size=None, t1=-1, t2=-1, channels=None):
    """Detection phase of ms3alg (p_mountainsort3.cpp:145-216, neighborhoodsorter.cpp:74-106).

    Returns (counts per central channel, times concatenated per channel in
    ascending order, absolute 0-based sample indices).
    """
    X = np.ascontiguousarray(X, np.float32)
    N, M = X.shape
    if channels is None:
        channels = np.arange(M, dtype=np.int32)
    channels = np.ascontiguousarray(channels, np.int32)
    t_start = 0 if t1 < 0 else int(t1)
    t_end = N - 1 if t2 < 0 else int(t2)
    Nr = t_end - t_start + 1
    cs = int(chunk_size) if chunk_size else detect_chunk_size(M, Nr, num_threads)
    counts = np.zeros(len(channels), np.int64)
    D = max(int(detect_interval), 0)
    cap = len(channels) * (Nr // (D + 1) + Nr // max(cs, 1) + 2)
    times = np.empty(max(cap, 1), np.int64)
    n = lib().orc_ms3alg_detect(_p(X), M, N, int(t1), int(t2), cs, DETECT_PADDING, _p(channels),
                                len(channels), float(detect_threshold), float(detect_interval),
                                int(detect_sign), _p(counts), _p(times), cap)
    if n < 0:
        raise RuntimeError("oracle capacity exceeded")
    return counts, times[:n].copy()


def firings_order(counts, times, channels=None):
    """(central channel, time) pairs sorted by time, ties in channel order:
    the stable time sort of p_mountainsort3.cpp:245-263 (get_sort_indices.cpp:52-61)."""
    counts = np.asarray(counts)
    if channels is None:
        channels = np.arange(len(counts))
    ch = np.repeat(np.asarray(channels), counts)
    order = np.argsort(times, kind="stable")
    return ch[order], np.asarray(times)[order]


# --------------------------------------------------------------------------
# a9, a11: neighborhoods and clips
# --------------------------------------------------------------------------
def channels_from_geom(geom, m, adjacency_radius):
    """get_channels_from_geom (p_mountainsort3.cpp:473-491).  geom: (M, D) coordinates
    (one CSV row per channel); m is 1-based.  Returns 1-based channel list."""
    geom = np.asarray(geom, np.float64)
    M = geom.shape[0]
    ret = [m]
    for m2 in range(1, M + 1):
        if m2 != m:
            s = 0.0
            for i in range(geom.shape[1]):
                tmp = geom[m - 1, i] - geom[m2 - 1, i]
                s += tmp * tmp
            dist = np.sqrt(s)
            if (adjacency_radius > 0 and dist <= adjacency_radius) or adjacency_radius < 0:
                ret.append(m2)
    return ret


def extract_clips(X, start_times, channels, T):
    """Clip gather: clip[w, t, i] = X[channels[w], start_i + t], 0 outside [0, N)
    (neighborhoodsorter.cpp:409-419; p_extract_clips.cpp:53-69).  channels are
    0-based.  Returns float32 (L, T, W)."""
    X = np.asarray(X)
    N, M = X.shape
    st = np.asarray(start_times, np.int64)
    ch = np.asarray(channels, np.int64)
    tt = st[:, None] + np.arange(T)[None, :]
    valid = (tt >= 0) & (tt < N)
    vals = X[np.clip(tt, 0, N - 1)][:, :, ch].astype(np.float32)
    vals[~valid] = 0
    return vals


def clip_tmid(T):
    """Tmid = (T+1)/2 - 1 (integer division), neighborhoodsorter.cpp:412."""
    return (int(T) + 1) // 2 - 1


def extract_clips_processor_starts(event_times, T):
    """ms3.extract_clips start index per event: (int64)(ET - Tmid) truncation
    (p_extract_clips.cpp:54,59)."""
    et = np.asarray(event_times, np.float64)
    return np.trunc(et - clip_tmid(T)).astype(np.int64)


# --------------------------------------------------------------------------
# a12-a15: PCA (packages/mountainsortalg/pca.cpp, neighborhoodsorter.cpp, sort_clips.cpp)
# --------------------------------------------------------------------------
def pca(X, K):
    """pca (pca.cpp:68-97) on an M x N float32 matrix given as (N, M).
    Returns (C (M, K) float32, F (N, K) float32, sigma (K,) float32)."""
    X = np.ascontiguousarray(X, np.float32)
    N, M = X.shape
    C = np.empty(M * K, np.float32)
    F = np.empty(max(K * N, 1), np.float32)
    s = np.empty(K, np.float32)
    lib().orc_pca(_p(X), M, N, K, _p(C), _p(F), _p(s))
    return C.reshape(K, M).T.copy(), F[:K * N].reshape(N, K), s


def pca_subsample_indices(N, max_samples):
    out = np.empty(int(max_samples) + 2 if N > max_samples else max(int(N), 1), np.int64)
    n = lib().orc_pca_subsample_indices(int(N), int(max_samples), _p(out))
    return out[:n].copy()


def pca_subsampled(X, K, max_samples=10000):
    """pca_subsampled (pca.cpp:99-127).  X given as (N, M)."""
    X = np.ascontiguousarray(X, np.float32)
    N, M = X.shape
    C = np.empty(M * K, np.float32)
    F = np.empty(max(K * N, 1), np.float32)
    s = np.empty(K, np.float32)
    lib().orc_pca_subsampled(_p(X), M, N, K, int(max_samples), _p(C), _p(F), _p(s))
    return C.reshape(K, M).T.copy(), F[:K * N].reshape(N, K), s


def dimension_reduce_clips(clips, nf=10, max_samples=10000):
    """Stage 1 (neighborhoodsorter.cpp:236-276).  clips (L, T, W) -> (L, nf, W)."""
    clips = np.ascontiguousarray(clips, np.float32)
    L, T, W = clips.shape
    out = np.empty((L, nf, W), np.float32)
    lib().orc_dimension_reduce_clips(_p(clips), W, T, L, nf, int(max_samples), _p(out))
    return out


def stage2_features(reduced, nf=10, max_samples=10000):
    """Stage 2 (sort_clips.cpp:58-70): reduced (L, nf1, W) viewed as (nf1*W) x L,
    pca_subsampled with num_features -> (L, nf) float32 features for isosplit5."""
    reduced = np.ascontiguousarray(reduced, np.float32)
    L = reduced.shape[0]
    X = reduced.reshape(L, -1)
    _, F, _ = pca_subsampled(X, nf, max_samples)
    return F


def pca_gram_fp64(X, K, max_samples=10000):
    """FP64 Gram-form restatement of pca_subsampled: G = X_s X_s^T, 10 power steps
    per component from sin(1..M), deflation of G, F = C^T X.  X given as (N, M).
    Differs from the float data form only by rounding."""
    X = np.asarray(X, np.float64)
    N, M = X.shape
    idx = pca_subsample_indices(N, max_samples)
    Xs = X[idx]
    G = Xs.T @ Xs
    C = np.zeros((M, K))
    sig = np.zeros(K)
    for k in range(K):
        v = np.sin(np.arange(1, M + 1, dtype=np.float64))
        v /= np.linalg.norm(v)
        s = 0.0
        for _ in range(10):
            v = G @ v
            s = np.linalg.norm(v)
            if s:
                v = v / s
        C[:, k] = v
        sig[k] = s
        P = np.eye(M) - np.outer(v, v)
        G = P @ G @ P
    return C, X @ C, sig


def dimension_reduce_clips_fp64(clips, nf=10, max_samples=10000):
    """Stage 1 in exact-product FP64 Gram form (pca_gram_fp64 per channel):
    the FP64 counterpart of dimension_reduce_clips.  (L, T, W) -> (L, nf, W)
    float64."""
    clips = np.asarray(clips)
    L, T, W = clips.shape
    out = np.zeros((L, nf, W), np.float64)
    for w in range(W):
        _, F, _ = pca_gram_fp64(clips[:, :, w], nf, max_samples)
        out[:, :, w] = F
    return out


def stage2_features_fp64(reduced, nf=10, max_samples=10000):
    """Stage 2 in FP64 Gram form on (L, nf1, W) reduced clips (sort_clips.cpp:58-70
    arithmetic in FP64): (L, nf) float64."""
    reduced = np.asarray(reduced, np.float64)
    L = reduced.shape[0]
    _, F, _ = pca_gram_fp64(reduced.reshape(L, -1), nf, max_samples)
    return F



# --------------------------------------------------------------------------
# Templates and the fit stage (packages/ms3/p_compute_templates.cpp,
# packages/mountainsortalg/globaltemplatecomputer.cpp, fitstagecomputer.cpp,
# fit_stage.cpp, p_mountainsort3.cpp:270-356)
# --------------------------------------------------------------------------
def time_chunk_infos(M, N, num_simultaneous, min_num_chunks):
    """get_time_chunk_infos (p_mountainsort3.cpp:507-533): [(t1, size, padding)]."""
    chunk = 1000000000 // (int(M) * int(num_simultaneous) * 4)
    if chunk > N:
        chunk = N
    if N < chunk * min_num_chunks:
        chunk = N // int(min_num_chunks)
    if chunk < 20000:
        chunk = 20000
    return [(t, min(chunk, N - t), 1000) for t in range(0, int(N), chunk)]


def compute_templates(X, times, labels, K, T):
    """ms3.compute_templates (p_compute_templates.cpp:24-118): X (N, M), times
    (L,) sample indices, labels (L,) 1..K.  Returns (templates (K, T, M)
    float32, i.e. the M x T x K .mda memory, counts (K,))."""
    X = np.ascontiguousarray(X, np.float32)
    N, M = X.shape
    t = np.ascontiguousarray(times, np.float64)
    lab = np.ascontiguousarray(labels, np.int32)
    out = np.empty((max(int(K), 0), int(T), M), np.float32)
    cnt = np.empty(max(int(K), 1), np.float64)
    lib().orc_compute_templates(_p(X), M, N, _p(t), _p(lab), len(t), int(K), int(T), _p(out),
                                _p(cnt))
    return out, cnt[:int(K)]


def global_template_events(times, N, M):
    """Events counted by GlobalTemplateComputer over get_time_chunk_infos(M, N, 1, 1)
    (globaltemplatecomputer.cpp:86-88, p_mountainsort3.cpp:270-289): an event at
    t0 in the chunk starting at t1 counts when padding <= t0 - t1 + padding <
    size, i.e. the last `padding` samples of each chunk's core are left out.
    Returns a boolean mask over the events."""
    times = np.asarray(times, np.int64)
    ok = np.zeros(len(times), bool)
    for t1, size, pad in time_chunk_infos(M, N, 1, 1):
        loc = times - t1 + pad
        ok |= (loc >= pad) & (loc < size)
    return ok


def global_templates(X, times, labels, K, T):
    """The ms3alg global templates: compute_templates on the events the chunk
    rule counts (one thread's summation order; the reference merges per-thread
    partial sums in completion order, so its last bits vary run to run)."""
    N, M = np.asarray(X).shape
    sel = global_template_events(times, N, M)
    return compute_templates(X, np.asarray(times)[sel], np.asarray(labels)[sel], K, T)


def fit_stage(Xc, times, labels, templates, thresh=0.1):
    """fit_stage (fit_stage.cpp:27-174) on one chunk Xc (N2, M), modified in
    place.  times: chunk-local (ascending), labels 1..K, templates (K, T, M).
    Returns (keep (L,) bool, passes)."""
    Xc = np.ascontiguousarray(Xc, np.float32)
    N2, M = Xc.shape
    tm = np.ascontiguousarray(templates, np.float32)
    K, T = tm.shape[0], tm.shape[1]
    t = np.ascontiguousarray(times, np.float64)
    lab = np.ascontiguousarray(labels, np.int32)
    keep = np.zeros(max(len(t), 1), np.uint8)
    passes = lib().orc_fit_stage(_p(Xc), M, N2, _p(t), _p(lab), len(t), _p(tm), K, T,
                                 float(thresh), _p(keep))
    return keep[:len(t)].astype(bool), int(passes)


def fit_stage_computer(X, times, labels, templates, num_threads, thresh=0.1):
    """FitStageComputer over get_time_chunk_infos(M, N, threads, threads)
    (p_mountainsort3.cpp:313-344, fitstagecomputer.cpp:55-104): every chunk
    padded by 1000 zero-filled samples fits its own copy with the events of
    the padded window; an event is used when its chunk-local time t1 satisfies
    padding <= t1 < size (fitstagecomputer.cpp:92-93).  Returns the sorted
    global indices of the used events."""
    X = np.asarray(X, np.float32)
    N, M = X.shape
    times = np.asarray(times, np.int64)
    labels = np.asarray(labels, np.int32)
    used = []
    for t1, size, pad in time_chunk_infos(M, N, num_threads, num_threads):
        N2 = size + 2 * pad
        Xc = np.zeros((N2, M), np.float32)
        a, b = t1 - pad, t1 + size + pad
        sa, sb = max(a, 0), min(b, N)
        Xc[sa - a:sb - a] = X[sa:sb]
        loc = times - t1 + pad
        sel = np.nonzero((loc >= 0) & (loc < N2))[0]
        keep, _ = fit_stage(Xc, loc[sel], labels[sel], templates, thresh)
        lt = loc[sel]
        ok = keep & (lt >= pad) & (lt < size)
        used.extend(sel[ok].tolist())
    return np.array(sorted(used), np.int64)


def sort_clips(clips, cluster, num_features=10, max_samples=10000):
    """sort_clips / sort_clips_subset (sort_clips.cpp:29-125) with the float-form
    PCA: the subset's clips viewed as (M*T) x L0 (:58-67), pca_subsampled
    (:70), the caller's clustering in place of isosplit5 (:73-88), and the
    branch method, depth first, labels of cluster k offset by the clusters
    found under clusters 1..k-1 (:97-123).  clips (L, T, M).  Returns (L,)
    labels."""
    clips = np.ascontiguousarray(clips, np.float32)
    L = clips.shape[0]

    def subset(indices):
        X = clips[indices].reshape(len(indices), -1)
        _, F, _ = pca_subsampled(X, num_features, max_samples)
        lab = np.asarray(cluster(F), dtype=np.int64).reshape(len(indices))
        K0 = int(lab.max()) if len(lab) else 0
        if K0 <= 1:
            return lab
        new = np.zeros(len(indices), np.int64)
        koff = 0
        for k in range(1, K0 + 1):
            sel = np.nonzero(lab == k)[0]
            if len(sel):
                sub = subset(indices[sel])
                new[sel] = koff + sub
                koff += int(sub.max())
        return new

    return subset(np.arange(L, dtype=np.int64))
