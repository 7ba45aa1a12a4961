"""The reference's CPU path, threaded as the reference threads it.

TEST INFRASTRUCTURE ONLY: this module times the CPU baseline that bench.py
reports beside the GPU figures.  It is never part of the product path.

Each function runs one stage of the front end with the reference's work
split and thread structure, on in-memory arrays in place of the reference's
.mda chunk I/O:

  bandpass_mt      p_bandpass_filter.cpp:170-229  one OpenMP iteration per
                   (chunk + 2*60000)-sample window: single-precision complex
                   DFT over all channels, gain x 1/n_fft, inverse DFT, core
                   written under a lock.  The DFT is pocketfft (scipy.fft,
                   complex64, one worker per window) in place of FFTW.
  whiten_mt        p_whiten.cpp:64-200  covariance and apply over 1e5-sample
                   chunks (OpenMP in oracle/msoracle.c), the whitening matrix
                   single-threaded (pca.cpp:171-193, 634-650).
  mask             p_mask_out_artifacts.cpp:24-104, single-threaded as there.
  ms3alg_clips_mt  p_mountainsort3.cpp:170-216 + neighborhoodsorter.cpp:74-110:
                   time chunks x neighborhood batches, detect + clip gather.
  pca_mt           p_mountainsort3.cpp:222-241 + neighborhoodsorter.cpp:153-160:
                   per neighborhood, dimension_reduce_clips and the first
                   sort_clips PCA.

Per element the arithmetic is that of oracle.core's sequential restatements;
the tests check the two against each other.
"""
import ctypes
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.fft as _sfft

from oracle import core as oc

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_bound = False


def _lib():
    global _bound
    L = oc.lib()
    if not _bound:
        L.orc_whiten_covariance_mt.argtypes = [_vp, _i64, _i64, _i64, ctypes.c_int, _vp]
        L.orc_whiten_covariance_mt.restype = None
        L.orc_whiten_apply_mt.argtypes = [_vp, _i64, _i64, _vp, _i64, ctypes.c_int, _vp]
        L.orc_whiten_apply_mt.restype = None
        L.orc_ms3alg_clips_mt.argtypes = [_vp, _i64, _i64, _i64, _i64, _vp, _vp, _i64, _i64,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_int, _i64,
                                          ctypes.c_int]
        L.orc_ms3alg_clips_mt.restype = _vp
        L.orc_clipset_counts.argtypes = [_vp, _vp]
        L.orc_clipset_counts.restype = None
        L.orc_clipset_get.argtypes = [_vp, _i64, _vp]
        L.orc_clipset_get.restype = None
        L.orc_clipset_free.argtypes = [_vp]
        L.orc_clipset_free.restype = None
        L.orc_pca_neighborhoods_mt.argtypes = [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64,
                                               ctypes.c_int, _vp, _vp]
        L.orc_pca_neighborhoods_mt.restype = None
        _bound = True
    return L


def bandpass_mt(X, samplerate, freq_min, freq_max, freq_wid=1000.0, num_threads=8,
                chunk_size=None, overlap=oc.BANDPASS_OVERLAP, windows=None):
    """ms3.bandpass_filter's threaded window loop.  X (N, M) float32.
    Returns (N, M) float32.  windows: restrict to these window start times
    (for timing a bounded sample); their cores are written, the rest of the
    output stays zero."""
    X = np.ascontiguousarray(X, np.float32)
    N, M = X.shape
    cs = int(chunk_size) if chunk_size else oc.bandpass_chunk_size(M, num_threads)
    n_fft = cs + 2 * overlap
    gain = (oc.define_kernel(n_fft, samplerate, freq_min, freq_max, freq_wid) *
            (1.0 / n_fft))[:, None]
    out = np.zeros((N, M), np.float32)
    lock = threading.Lock()
    starts = list(range(0, N, cs)) if windows is None else list(windows)

    def one(t):
        win = np.zeros((n_fft, M), np.complex64)
        with lock:   # readChunk under critical(lock1), :182-188
            a = t - overlap
            sa, sb = max(a, 0), min(a + n_fft, N)
            if sb > sa:
                win[sa - a:sb - a].real = X[sa:sb]
        Y = _sfft.fft(win, axis=0, workers=1, overwrite_x=True)
        for a in range(0, n_fft, 8192):     # data_out *= kernel0[i]*factor, :88-95
            Y[a:a + 8192] = Y[a:a + 8192] * gain[a:a + 8192]
        y = _sfft.ifft(Y, axis=0, workers=1, norm="forward", overwrite_x=True).real
        n = min(cs, N - t)
        with lock:   # writeChunk under critical(lock1), :198-214
            out[t:t + n] = y[overlap:overlap + n]

    with ThreadPoolExecutor(max_workers=int(num_threads)) as ex:
        list(ex.map(one, starts))
    return out


def whiten_covariance_mt(X, num_threads=8, chunk_size=100000):
    X = np.ascontiguousarray(X, np.float32)
    N, M = X.shape
    C = np.empty((M, M), np.float64)
    _lib().orc_whiten_covariance_mt(oc._p(X), M, N, int(chunk_size), int(num_threads), oc._p(C))
    return C


def whiten_apply_mt(X, W, num_threads=8, chunk_size=100000):
    """The apply pass with the 1e-4 quantize (p_whiten.cpp:160-200).  W is
    symmetric in the reference, so the row-major W here is its Mda memory."""
    X = np.ascontiguousarray(X, np.float32)
    N, M = X.shape
    Wc = np.ascontiguousarray(W, np.float64)
    Y = np.empty((N, M), np.float32)
    _lib().orc_whiten_apply_mt(oc._p(X), M, N, oc._p(Wc), int(chunk_size), int(num_threads),
                               oc._p(Y))
    return Y


def whiten_mt(X, num_threads=8, chunk_size=100000):
    """ms3.whiten: covariance, whitening matrix, apply.  Returns (Y, W)."""
    C = whiten_covariance_mt(X, num_threads, chunk_size)
    W = oc.whitening_matrix(C)
    return whiten_apply_mt(X, W, num_threads, chunk_size), W


def _flat_neighborhoods(neighborhoods):
    ch = np.ascontiguousarray(np.concatenate([np.asarray(n, np.int32) for n in neighborhoods])
                              if len(neighborhoods) else np.zeros(0, np.int32), np.int32)
    off = np.ascontiguousarray(np.concatenate([[0], np.cumsum([len(n) for n in neighborhoods])]),
                               np.int64)
    return ch, off


def ms3alg_clips_mt(X, neighborhoods, detect_threshold=3.0, detect_interval=10, detect_sign=-1,
                    clip_size=50, num_threads=8, chunk_size=None, padding=1000):
    """The ms3alg read phase.  X (N, M) float32; neighborhoods: lists of
    0-based channels, central channel first.  The time chunks follow
    get_time_chunk_infos(M, N, number of batches, 1) unless chunk_size is
    given.  Returns a list of (L_g, T, W_g) float32 clip arrays."""
    X = np.ascontiguousarray(X, np.float32)
    N, M = X.shape
    G = len(neighborhoods)
    batch = max(1, min(int(num_threads), M))                    # :171
    if not chunk_size:
        nb = (G + batch - 1) // batch
        chunk_size = oc.time_chunk_infos(M, N, nb, 1)[0][1]     # :191
    ch, off = _flat_neighborhoods(neighborhoods)
    L = _lib()
    cs = L.orc_ms3alg_clips_mt(oc._p(X), M, N, int(chunk_size), int(padding), oc._p(ch),
                               oc._p(off), G, batch, float(detect_threshold),
                               float(detect_interval), int(detect_sign), int(clip_size),
                               int(num_threads))
    try:
        cnt = np.zeros(max(G, 1), np.int64)
        L.orc_clipset_counts(cs, oc._p(cnt))
        out = []
        for g in range(G):
            c = np.empty((int(cnt[g]), clip_size, len(neighborhoods[g])), np.float32)
            if c.size:
                L.orc_clipset_get(cs, g, oc._p(c))
            out.append(c)
        return out
    finally:
        L.orc_clipset_free(cs)


def pca_mt(clips, num_features=10, num_features2=10, max_samples=10000, num_threads=8,
           return_results=False):
    """Per-neighborhood stage-1 and stage-2 PCA, neighborhoods in batches of
    num_threads.  clips: list of (L_g, T, W_g) float32 arrays."""
    clips = [np.ascontiguousarray(c, np.float32) for c in clips]
    G = len(clips)
    T = clips[0].shape[1] if G else 0
    W = np.array([c.shape[2] for c in clips] or [0], np.int64)
    Ls = np.array([c.shape[0] for c in clips] or [0], np.int64)
    ptrs = (ctypes.c_void_p * max(G, 1))(*[c.ctypes.data for c in clips])
    red = feat = None
    rp = fp = None
    if return_results:
        red = [np.empty((c.shape[0], num_features, c.shape[2]), np.float32) for c in clips]
        feat = [np.empty((c.shape[0], num_features2), np.float32) for c in clips]
        rp = (ctypes.c_void_p * max(G, 1))(*[r.ctypes.data for r in red])
        fp = (ctypes.c_void_p * max(G, 1))(*[f.ctypes.data for f in feat])
    batch = max(1, int(num_threads))
    _lib().orc_pca_neighborhoods_mt(ctypes.cast(ptrs, _vp), oc._p(W), oc._p(Ls), G, T,
                                    int(num_features), int(num_features2), int(max_samples),
                                    batch, int(num_threads),
                                    ctypes.cast(rp, _vp) if rp is not None else None,
                                    ctypes.cast(fp, _vp) if fp is not None else None)
    return (red, feat) if return_results else None
