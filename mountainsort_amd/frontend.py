"""Device-resident operators of the MountainSort front end over the C ABI.

These mirror the reference's processors and in-process stages with the same
parameter names (packages/ms3/mountainsort3_main.cpp:145-198,343-350;
packages/mountainsortalg/mountainsortalg_main.cpp:31-46) but take and return
torch tensors that already live on the GPU.  torch only provides device memory
and the current stream; all arithmetic runs in the HIP kernels of libmsamd.so.

Layouts (identical to the .mda memory order of the reference):
  timeseries M x N  -> tensor (N, M), element (m, t) at m + M*t
  clips W x T x L   -> tensor (L, T, W)
  M x M matrices    -> tensor (M, M) indexed [row, col]; the C ABI receives
                       the column-major (Mda) buffer.
"""
import ctypes

import torch

from . import _capi

DETECT_PADDING = 1000      # p_mountainsort3.cpp:515
BANDPASS_OVERLAP = 60000   # p_bandpass_filter.cpp:150


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _check_ts(X, name="timeseries", dtype=torch.float32):
    if not isinstance(X, torch.Tensor) or not X.is_cuda:
        raise TypeError(f"{name} must be a CUDA tensor")
    if X.dim() != 2 or X.dtype != dtype or not X.is_contiguous():
        raise ValueError(f"{name} must be a contiguous ({dtype}) tensor of shape (N, M)")


def _channel_list(channels, M, dev):
    ch = torch.as_tensor(channels, dtype=torch.int32).flatten()
    if ch.numel() and (int(ch.min()) < 0 or int(ch.max()) >= M):
        raise ValueError(f"channel index outside [0, {M})")
    return ch.to(dev).contiguous()


def _workspace(nbytes, device):
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)


# --------------------------------------------------------------------------
# bandpass filter (a1-a3)
# --------------------------------------------------------------------------
def bandpass_chunk_size(M, num_threads):
    """Core chunk of the filter windows (p_bandpass_filter.cpp:147-155)."""
    return int(_capi.lib().msamd_bandpass_chunk_size(int(M), int(num_threads)))


def bandpass_filter(X, samplerate, freq_min, freq_max, freq_wid=1000.0, quantization_unit=0.0,
                    subsample_factor=1, num_threads=8, chunk_size=None, overlap=BANDPASS_OVERLAP,
                    out=None):
    """ms3.bandpass_filter on a device-resident (N, M) float32 or int16 tensor.

    Windows follow the reference's chunk plan for `num_threads`
    (p_bandpass_filter.cpp:147-155) unless chunk_size is given.  Returns
    float32 (N/k, M), or int16 when quantization_unit > 0.  freq_max == 0 is
    the reference's file copy (:117-119) and returns a copy here.
    """
    if not isinstance(X, torch.Tensor) or not X.is_cuda or X.dim() != 2 or not X.is_contiguous():
        raise ValueError("timeseries must be a contiguous CUDA tensor of shape (N, M)")
    if freq_max == 0:
        return X.clone()
    dtypes = {torch.float32: -3, torch.int16: -4}
    if X.dtype not in dtypes:
        raise ValueError("timeseries dtype must be float32 or int16")
    N, M = X.shape
    dev = X.device
    k = int(subsample_factor) if subsample_factor else 1
    cs = int(chunk_size) if chunk_size else bandpass_chunk_size(M, num_threads)
    L = _capi.lib()
    nb = L.msamd_bandpass_workspace_size(M, N, cs, int(overlap))
    if nb < 0:
        raise _capi.MsamdError("msamd_bandpass_workspace_size", 5,
                               f"FFT length {cs + 2 * int(overlap)} not supported")
    ws = _workspace(nb, dev)
    q = float(quantization_unit)
    if out is None:
        out = torch.empty((N // k, M), dtype=torch.int16 if q else torch.float32, device=dev)
    _capi.call("msamd_bandpass_filter", _ptr(X), dtypes[X.dtype], M, N, float(samplerate),
               float(freq_min), float(freq_max), float(freq_wid), cs, int(overlap), k, q,
               _ptr(out), _ptr(ws), ws.numel(), _stream(dev))
    return out


# --------------------------------------------------------------------------
# detection (a8-a10)
# --------------------------------------------------------------------------
def detect_chunk_size(M, N, num_threads):
    """Chunk length of the ms3alg time-chunk plan (p_mountainsort3.cpp:170-172,507-533)."""
    return int(_capi.lib().msamd_detect_chunk_size(int(M), int(N), int(num_threads)))


def detect_events(X, detect_threshold=3.0, detect_interval=10, detect_sign=0, num_threads=8,
                  chunk_size=None, t1=-1, t2=-1, channels=None, padding=DETECT_PADDING):
    """Detection phase of mountainsortalg.ms3alg on every central channel.

    Returns (offsets, times): int64 device tensors; the events of
    channels[c] are times[offsets[c]:offsets[c+1]] (absolute 0-based sample
    indices, ascending).  num_threads is the reference thread count that
    fixes the chunk plan (p_mountainsort3.cpp:170-172).
    """
    _check_ts(X)
    N, M = X.shape
    dev = X.device
    if channels is None:
        channels = torch.arange(M, dtype=torch.int32, device=dev)
    else:
        channels = _channel_list(channels, M, dev)
    Mc = int(channels.numel())
    t_start = 0 if t1 < 0 else int(t1)
    t_end = N - 1 if t2 < 0 else int(t2)
    n_range = t_end - t_start + 1
    cs = int(chunk_size) if chunk_size else detect_chunk_size(M, n_range, num_threads)
    L = _capi.lib()
    nb = L.msamd_detect_workspace_size(n_range, cs, Mc, float(detect_interval))
    ws = _workspace(nb, dev)
    total = ctypes.c_int64(0)
    st = _stream(dev)
    _capi.call("msamd_detect_events", _ptr(X), M, N, int(t1), int(t2), cs, int(padding),
               _ptr(channels), Mc, float(detect_threshold), float(detect_interval),
               int(detect_sign), _ptr(ws), ws.numel(), ctypes.byref(total), st)
    times = torch.empty(max(total.value, 1), dtype=torch.int64, device=dev)
    offsets = torch.empty(Mc + 1, dtype=torch.int64, device=dev)
    _capi.call("msamd_detect_fetch", _ptr(ws), Mc, _ptr(times), _ptr(offsets), st)
    return offsets, times[:total.value]


# --------------------------------------------------------------------------
# clips (a11)
# --------------------------------------------------------------------------
def clip_tmid(T):
    return (int(T) + 1) // 2 - 1


def extract_clips(X, starts, channels, clip_size):
    """Gather clips (L, T, W) with clip[i, t, w] = X[starts[i] + t, channels[w]]
    (zero outside the recording).  starts: int64 first-sample indices."""
    _check_ts(X)
    N, M = X.shape
    dev = X.device
    starts = torch.as_tensor(starts, dtype=torch.int64).to(dev).contiguous()
    channels = _channel_list(channels, M, dev)
    L, W, T = int(starts.numel()), int(channels.numel()), int(clip_size)
    clips = torch.empty((L, T, W), dtype=torch.float32, device=dev)
    _capi.call("msamd_extract_clips", _ptr(X), M, N, _ptr(starts), L, _ptr(channels), W, T,
               _ptr(clips), _stream(dev))
    return clips


def extract_clips_batched(X, times, event_offsets, nbhd_channels, nbhd_offsets, clip_size):
    """Clips of G neighborhoods in one launch.  Returns (clips_flat, clip_offsets):
    neighborhood g's (L_g, T, W_g) clips start at clips_flat[clip_offsets[g]]."""
    _check_ts(X)
    N, M = X.shape
    dev = X.device
    T = int(clip_size)
    eo = torch.as_tensor(event_offsets, dtype=torch.int64).to(dev).contiguous()
    no = torch.as_tensor(nbhd_offsets, dtype=torch.int64).to(dev).contiguous()
    nc = _channel_list(nbhd_channels, M, dev)
    times = times.to(dev).contiguous()
    G = int(no.numel()) - 1
    sizes = (eo[1:] - eo[:-1]) * (no[1:] - no[:-1]) * T
    co = torch.zeros(G + 1, dtype=torch.int64, device=dev)
    co[1:] = torch.cumsum(sizes, 0)
    total_elems = int(co[-1].item())
    clips = torch.empty(max(total_elems, 1), dtype=torch.float32, device=dev)
    total_events = int(eo[-1].item())
    _capi.call("msamd_extract_clips_batched", _ptr(X), M, N, _ptr(times), _ptr(eo), total_events,
               _ptr(nc), _ptr(no), G, T, _ptr(co), _ptr(clips), _stream(dev))
    return clips[:total_elems], co


# --------------------------------------------------------------------------
# PCA features (a12-a15)
# --------------------------------------------------------------------------
def pca_features(clips_flat, clip_offsets, event_offsets, nbhd_offsets, clip_size,
                 num_features_per_channel=10, num_features=10, max_pca_samples=10000,
                 tiles=None):
    """Two-stage PCA features of every neighborhood (neighborhoodsorter.cpp:236-276,
    sort_clips.cpp:58-70).

    clips_flat/clip_offsets as returned by extract_clips_batched; event_offsets
    and nbhd_offsets are the (G+1) CSR offsets of events and channels.  Returns
    (reduced_flat, reduced_offsets, features_flat, feature_offsets): per
    neighborhood g, reduced (L, nf1, W) at reduced_flat[reduced_offsets[g]:] and
    features (L, nf2) at features_flat[feature_offsets[g]:].  tiles: the
    plan's tile sizes (see pca_autotune); None uses the defaults.
    """
    import numpy as np
    dev = clips_flat.device
    eo = torch.as_tensor(event_offsets).cpu().numpy().astype(np.int64)
    no = torch.as_tensor(nbhd_offsets).cpu().numpy().astype(np.int64)
    L = np.ascontiguousarray(eo[1:] - eo[:-1], dtype=np.int64)
    W = np.ascontiguousarray(no[1:] - no[:-1], dtype=np.int32)
    if tiles is not None:
        plan = PcaPlan(L, W, clip_size, num_features_per_channel, num_features, max_pca_samples,
                       tiles=tiles, device=dev)
        out = plan.run(clips_flat)
        plan.close()  # the outputs stay alive; the kernels are enqueued on the stream
        return out
    G = len(L)
    T, nf1, nf2 = int(clip_size), int(num_features_per_channel), int(num_features)
    lib = _capi.lib()
    pL = L.ctypes.data_as(ctypes.c_void_p)
    pW = W.ctypes.data_as(ctypes.c_void_p)
    nb = lib.msamd_pca_workspace_size(G, pL, pW, T, nf1, nf2, int(max_pca_samples))
    if nb < 0:
        raise _capi.MsamdError("msamd_pca_workspace_size", 1, "unsupported neighborhood shapes")
    ws = _workspace(nb, dev)
    red_sizes = W.astype(np.int64) * nf1 * L
    red_off = np.concatenate([[0], np.cumsum(red_sizes)])
    feat_off = np.concatenate([[0], np.cumsum(nf2 * L)])
    reduced = torch.empty(max(int(red_off[-1]), 1), dtype=torch.float32, device=dev)
    features = torch.empty(max(int(feat_off[-1]), 1), dtype=torch.float32, device=dev)
    _capi.call("msamd_pca_features", _ptr(clips_flat), G, pL, pW, T, nf1, nf2,
               int(max_pca_samples), _ptr(reduced), _ptr(features), _ptr(ws), ws.numel(),
               _stream(dev))
    return reduced, red_off, features, feat_off


def _pca_shapes(num_events, widths):
    import numpy as np
    L = np.ascontiguousarray(np.asarray(num_events, dtype=np.int64).ravel())
    W = np.ascontiguousarray(np.asarray(widths, dtype=np.int32).ravel())
    if L.shape != W.shape or L.size == 0:
        raise ValueError("num_events and widths must be non-empty and of one length")
    return L, W


class PcaPlan:
    """A reusable batched two-stage PCA over fixed neighborhood shapes
    (msamd_pca_plan_*): the host tables are built and uploaded once, and
    run() only enqueues the batched kernels -- or, with graph=True, replays
    them as one captured HIP graph.

    num_events[g], widths[g]: events and channels of neighborhood g; clips
    are packed as extract_clips_batched returns them.  tiles: (Gram columns
    per item, stage-1 events per item, stage-2 events per item) or None for
    the defaults (see pca_autotune).  Outputs live in the plan (reduced,
    features) and are overwritten by every run.
    """

    def __init__(self, num_events, widths, clip_size, num_features_per_channel=10,
                 num_features=10, max_pca_samples=10000, tiles=None, device=None):
        import numpy as np
        self.device = torch.device(device) if device is not None else torch.device("cuda", 0)
        L, W = _pca_shapes(num_events, widths)
        self.G = int(L.size)
        self.T, self.nf1, self.nf2 = int(clip_size), int(num_features_per_channel), int(num_features)
        self.max_s = int(max_pca_samples)
        self._tiles_arg = None
        if tiles is not None:
            self._tiles_arg = np.ascontiguousarray(np.asarray(tiles, dtype=np.int32).ravel())
            if self._tiles_arg.size != 3:
                raise ValueError("tiles must hold three integers")
        pt = self._tiles_arg.ctypes.data_as(ctypes.c_void_p) if self._tiles_arg is not None else None
        pL, pW = L.ctypes.data_as(ctypes.c_void_p), W.ctypes.data_as(ctypes.c_void_p)
        lib = _capi.lib()
        # host tables built once (deferred create), then uploaded into the workspace
        self._plan = ctypes.c_void_p()
        _capi.call("msamd_pca_plan_create", self.G, pL, pW, self.T, self.nf1, self.nf2,
                   self.max_s, pt, None, 0, _stream(self.device), ctypes.addressof(self._plan))
        nb = lib.msamd_pca_plan_workspace_bytes(self._plan)
        self._ws = _workspace(nb, self.device)
        self.clip_elems = int((W.astype(np.int64) * self.T * L).sum())
        self.red_off = np.concatenate([[0], np.cumsum(W.astype(np.int64) * self.nf1 * L)])
        self.feat_off = np.concatenate([[0], np.cumsum(self.nf2 * L)])
        self.reduced = torch.empty(max(int(self.red_off[-1]), 1), dtype=torch.float32,
                                   device=self.device)
        self.features = torch.empty(max(int(self.feat_off[-1]), 1), dtype=torch.float32,
                                    device=self.device)
        _capi.call("msamd_pca_plan_attach", self._plan, _ptr(self._ws), self._ws.numel(),
                   _stream(self.device))

    @property
    def tiles(self):
        import numpy as np
        t = np.zeros(3, dtype=np.int32)
        _capi.lib().msamd_pca_plan_tiles(self._plan, t.ctypes.data_as(ctypes.c_void_p))
        return tuple(int(x) for x in t)

    def run(self, clips_flat, graph=False):
        """Enqueue the PCA of clips_flat; returns (reduced, red_off, features, feat_off)."""
        if self._plan is None or not self._plan.value:
            raise RuntimeError("PcaPlan is closed")
        if (clips_flat.dtype != torch.float32 or clips_flat.device != self.device
                or not clips_flat.is_contiguous() or clips_flat.numel() < self.clip_elems):
            raise ValueError("clips must be a contiguous float32 tensor of the plan's shapes "
                             "on the plan's device")
        fn = "msamd_pca_plan_launch_graph" if graph else "msamd_pca_plan_run"
        _capi.call(fn, self._plan, _ptr(clips_flat), _ptr(self.reduced), _ptr(self.features),
                   _stream(self.device))
        return self.reduced, self.red_off, self.features, self.feat_off

    def run_timeseries(self, X, times, nbhd_channels, nbhd_offsets):
        """The PCA with the clip gather fused into the stage-1 kernels
        (msamd_pca_plan_run_timeseries): clips are read from the timeseries
        X (N, M) at times - Tmid, no clip array is made.  times: the event
        centres packed neighborhood after neighborhood (the plan's
        num_events); nbhd_channels/nbhd_offsets as for extract_clips_batched.
        Returns (reduced, red_off, features, feat_off), bit-identical to
        run() on the gathered clips."""
        if self._plan is None or not self._plan.value:
            raise RuntimeError("PcaPlan is closed")
        _check_ts(X)
        if X.device != self.device:
            raise ValueError("X must be on the plan's device")
        N, M = X.shape
        times = times.to(self.device, torch.int64).contiguous()
        no = torch.as_tensor(nbhd_offsets, dtype=torch.int64).to(self.device).contiguous()
        nc = _channel_list(nbhd_channels, M, self.device)
        if no.numel() != self.G + 1:
            raise ValueError("nbhd_offsets must hold G + 1 offsets")
        self._keep = (times, no, nc)  # alive until the next run (kernels are enqueued)
        _capi.call("msamd_pca_plan_run_timeseries", self._plan, _ptr(X), M, N, _ptr(times),
                   _ptr(nc), _ptr(no), _ptr(self.reduced), _ptr(self.features),
                   _stream(self.device))
        return self.reduced, self.red_off, self.features, self.feat_off

    def close(self):
        if getattr(self, "_plan", None) is not None and self._plan.value:
            _capi.lib().msamd_pca_plan_destroy(self._plan)
        self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pca_autotune(clips_flat, num_events, widths, clip_size, num_features_per_channel=10,
                 num_features=10, max_pca_samples=10000):
    """Choose the PCA plan's tiles for this workload by timing candidates on
    the device (msamd_pca_autotune).  Returns (tiles, milliseconds)."""
    import numpy as np
    dev = clips_flat.device
    L, W = _pca_shapes(num_events, widths)
    G, T = int(L.size), int(clip_size)
    nf1, nf2 = int(num_features_per_channel), int(num_features)
    pL, pW = L.ctypes.data_as(ctypes.c_void_p), W.ctypes.data_as(ctypes.c_void_p)
    lib = _capi.lib()
    nb = lib.msamd_pca_autotune_workspace_size(G, pL, pW, T, nf1, nf2, int(max_pca_samples))
    if nb < 0:
        raise _capi.MsamdError("msamd_pca_autotune_workspace_size", 1, "unsupported shapes")
    ws = _workspace(nb, dev)
    red = torch.empty(max(int((W.astype(np.int64) * nf1 * L).sum()), 1), dtype=torch.float32,
                      device=dev)
    feat = torch.empty(max(int((nf2 * L).sum()), 1), dtype=torch.float32, device=dev)
    tiles = np.zeros(3, dtype=np.int32)
    ms = ctypes.c_float()
    _capi.call("msamd_pca_autotune", _ptr(clips_flat), G, pL, pW, T, nf1, nf2,
               int(max_pca_samples), _ptr(red), _ptr(feat), _ptr(ws), ws.numel(),
               _stream(dev), tiles.ctypes.data_as(ctypes.c_void_p), ctypes.addressof(ms))
    return tuple(int(x) for x in tiles), float(ms.value)


def pca_features_from_timeseries(X, times, event_offsets, nbhd_channels, nbhd_offsets,
                                 clip_size, num_features_per_channel=10, num_features=10,
                                 max_pca_samples=10000, tiles=None):
    """pca_features with the clip gather fused into the stage-1 kernels: the
    same (reduced, red_off, features, feat_off), bit for bit, as
    pca_features(extract_clips_batched(X, times, ...)), without the clip
    array.  times: event centres, neighborhood g's at
    times[event_offsets[g]:event_offsets[g+1]]."""
    import numpy as np
    eo = torch.as_tensor(event_offsets).cpu().numpy().astype(np.int64)
    no = torch.as_tensor(nbhd_offsets).cpu().numpy().astype(np.int64)
    L = np.ascontiguousarray(eo[1:] - eo[:-1], dtype=np.int64)
    W = np.ascontiguousarray(no[1:] - no[:-1], dtype=np.int32)
    plan = PcaPlan(L, W, clip_size, num_features_per_channel, num_features, max_pca_samples,
                   tiles=tiles, device=X.device)
    out = plan.run_timeseries(X, times[int(eo[0]):int(eo[-1])], nbhd_channels, nbhd_offsets)
    plan.close()  # the outputs stay alive; the kernels are enqueued on the stream
    return out


def pca_autotune_timeseries(X, times, event_offsets, nbhd_channels, nbhd_offsets, clip_size,
                            num_features_per_channel=10, num_features=10,
                            max_pca_samples=10000):
    """pca_autotune for the fused form (msamd_pca_autotune_timeseries).
    Returns (tiles, milliseconds)."""
    import numpy as np
    _check_ts(X)
    N, M = X.shape
    dev = X.device
    eo = torch.as_tensor(event_offsets).cpu().numpy().astype(np.int64)
    no_h = torch.as_tensor(nbhd_offsets).cpu().numpy().astype(np.int64)
    L, W = _pca_shapes(eo[1:] - eo[:-1], no_h[1:] - no_h[:-1])
    G, T = int(L.size), int(clip_size)
    nf1, nf2 = int(num_features_per_channel), int(num_features)
    pL, pW = L.ctypes.data_as(ctypes.c_void_p), W.ctypes.data_as(ctypes.c_void_p)
    lib = _capi.lib()
    nb = lib.msamd_pca_autotune_workspace_size(G, pL, pW, T, nf1, nf2, int(max_pca_samples))
    if nb < 0:
        raise _capi.MsamdError("msamd_pca_autotune_workspace_size", 1, "unsupported shapes")
    ws = _workspace(nb, dev)
    red = torch.empty(max(int((W.astype(np.int64) * nf1 * L).sum()), 1), dtype=torch.float32,
                      device=dev)
    feat = torch.empty(max(int((nf2 * L).sum()), 1), dtype=torch.float32, device=dev)
    t = times[int(eo[0]):int(eo[-1])].to(dev, torch.int64).contiguous()
    no = torch.as_tensor(nbhd_offsets, dtype=torch.int64).to(dev).contiguous()
    nc = _channel_list(nbhd_channels, M, dev)
    tiles = np.zeros(3, dtype=np.int32)
    ms = ctypes.c_float()
    _capi.call("msamd_pca_autotune_timeseries", _ptr(X), M, N, _ptr(t), _ptr(nc), _ptr(no), G,
               pL, pW, T, nf1, nf2, int(max_pca_samples), _ptr(red), _ptr(feat), _ptr(ws),
               ws.numel(), _stream(dev), tiles.ctypes.data_as(ctypes.c_void_p),
               ctypes.addressof(ms))
    return tuple(int(x) for x in tiles), float(ms.value)


# --------------------------------------------------------------------------
# whitening (a4-a6)
# --------------------------------------------------------------------------
def _row_keep(row_keep, N, dev):
    if row_keep is None:
        return None
    if (row_keep.dtype != torch.int32 or row_keep.device != dev or not row_keep.is_contiguous()
            or row_keep.numel() < (N + 31) // 32):
        raise ValueError("row_keep must be mask_rows' int32 bitmap on the timeseries' device")
    return row_keep


def whiten_covariance(X, row_keep=None):
    """X X^T / (N-1) as an (M, M) float64 tensor (p_whiten.cpp:64-118).

    row_keep: mask_rows' bitmap; rows whose bit is clear are read as zero
    (the covariance of the masked timeseries without writing it)."""
    _check_ts(X)
    N, M = X.shape
    dev = X.device
    L = _capi.lib()
    ws = _workspace(L.msamd_whiten_covariance_workspace_size(M, N), dev)
    C = torch.empty((M, M), dtype=torch.float64, device=dev)
    rk = _row_keep(row_keep, N, dev)
    if rk is None:
        _capi.call("msamd_whiten_covariance", _ptr(X), M, N, _ptr(C), _ptr(ws), ws.numel(),
                   _stream(dev))
    else:
        _capi.call("msamd_whiten_covariance_rows", _ptr(X), M, N, _ptr(rk), _ptr(C), _ptr(ws),
                   ws.numel(), _stream(dev))
    return C


def whitening_matrix(C, exact=False, return_info=False, fallback=None):
    """W from the covariance.

    exact=False: the reference's deflated power method (pca.cpp:634-650).
    exact=True: W = C^{-1/2}, the symmetric inverse square root the power
    method approximates (coupled Newton-Schulz in FP64; zero rows and columns
    for channels with a zero diagonal).  With return_info the exact mode also
    returns (iterations, residual).

    The power method's workgroups exchange values and must run together; if
    other work on the device keeps them apart its waits give up
    (MsamdError with code ERR_TIMEOUT).  fallback="exact" then warns and
    returns the exact form instead (a different W: the power method's is not
    converged)."""
    if not exact and fallback not in (None, "exact"):
        raise ValueError("fallback must be None or 'exact'")
    if not exact and fallback == "exact":
        try:
            return whitening_matrix(C, exact=False, return_info=return_info)
        except _capi.MsamdError as e:
            if e.code != _capi.ERR_TIMEOUT:
                raise
            import warnings
            warnings.warn(f"{e}; using the exact whitening matrix", RuntimeWarning)
            return whitening_matrix(C, exact=True, return_info=return_info)
    M = C.shape[0]
    dev = C.device
    Ccm = C.to(torch.float64).t().contiguous()
    L = _capi.lib()
    Wcm = torch.empty((M, M), dtype=torch.float64, device=dev)
    if exact:
        ws = _workspace(L.msamd_whitening_matrix_exact_workspace_size(M), dev)
        iters, resid = ctypes.c_int32(), ctypes.c_double()
        _capi.call("msamd_whitening_matrix_exact", _ptr(Ccm), M, _ptr(Wcm), _ptr(ws),
                   ws.numel(), _stream(dev), ctypes.addressof(iters), ctypes.addressof(resid))
        W = Wcm.t()
        return (W, (int(iters.value), float(resid.value))) if return_info else W
    ws = _workspace(L.msamd_whitening_matrix_workspace_size(M), dev)
    _capi.call("msamd_whitening_matrix", _ptr(Ccm), M, _ptr(Wcm), _ptr(ws), ws.numel(),
               _stream(dev))
    return Wcm.t()  # [row, col] view of the column-major result


def whiten_apply(X, W, quantization_unit=0.0, out=None, row_keep=None, xxt=None):
    """quantize(W applied to every sample, 1e-4) (p_whiten.cpp:153-206).

    Returns float32 (N, M); with quantization_unit > 0 returns int16 values
    (int)(y/q + 0.5) (p_whiten.cpp:445-451).  row_keep: mask_rows' bitmap;
    masked rows are read as zero.  xxt: the covariance of these rows (as
    whiten_covariance returned it); its diagonal bounds the samples, which
    spares the library a pass over X to find their range."""
    _check_ts(X)
    N, M = X.shape
    dev = X.device
    Wcm = W.to(device=dev, dtype=torch.float64).t().contiguous()
    L = _capi.lib()
    ws = _workspace(L.msamd_whiten_apply_workspace_size(M), dev)
    q = float(quantization_unit)
    rk = _row_keep(row_keep, N, dev)

    C = None
    if xxt is not None:
        if xxt.shape != (M, M) or xxt.device != dev:
            raise ValueError("xxt must be the (M, M) covariance on the timeseries' device")
        C = xxt.to(torch.float64).contiguous()  # symmetric: row- and column-major agree

    def call(Yp, qq, Yqp):
        if C is not None:
            _capi.call("msamd_whiten_apply_cov", _ptr(X), M, N, _ptr(Wcm), _ptr(C),
                       _ptr(rk) if rk is not None else None, Yp, qq, Yqp, _ptr(ws), ws.numel(),
                       _stream(dev))
        elif rk is None:
            _capi.call("msamd_whiten_apply", _ptr(X), M, N, _ptr(Wcm), Yp, qq, Yqp, _ptr(ws),
                       ws.numel(), _stream(dev))
        else:
            _capi.call("msamd_whiten_apply_rows", _ptr(X), M, N, _ptr(Wcm), _ptr(rk), Yp, qq,
                       Yqp, _ptr(ws), ws.numel(), _stream(dev))

    if q > 0:
        Yq = torch.empty((N, M), dtype=torch.int16, device=dev)
        call(None, q, _ptr(Yq))
        return Yq
    Y = out if out is not None else torch.empty((N, M), dtype=torch.float32, device=dev)
    call(_ptr(Y), 0.0, None)
    return Y


def whiten_clips(clips, W, quantization_unit=0.0):
    """ms3.whiten_clips (p_whiten.cpp:295-345) on device clips of shape (L, T, M)
    (the M x T x L .mda layout): W applied to every time sample, no 1e-4
    quantize.  Returns float32 (L, T, M), or int16 (int)(y/q + 0.5) values
    when quantization_unit > 0."""
    if (not isinstance(clips, torch.Tensor) or not clips.is_cuda or clips.dim() != 3
            or clips.dtype != torch.float32 or not clips.is_contiguous()):
        raise ValueError("clips must be a contiguous float32 CUDA tensor of shape (L, T, M)")
    L, T, M = clips.shape
    dev = clips.device
    q = float(quantization_unit)
    if L * T == 0:
        return torch.empty((L, T, M), dtype=torch.int16 if q > 0 else torch.float32, device=dev)
    Wcm = W.to(device=dev, dtype=torch.float64).t().contiguous()
    lib = _capi.lib()
    ws = _workspace(lib.msamd_whiten_apply_workspace_size(M), dev)
    if q > 0:
        out = torch.empty((L, T, M), dtype=torch.int16, device=dev)
        _capi.call("msamd_whiten_clips", _ptr(clips), M, T, L, _ptr(Wcm), None, q, _ptr(out),
                   _ptr(ws), ws.numel(), _stream(dev))
        return out
    out = torch.empty((L, T, M), dtype=torch.float32, device=dev)
    _capi.call("msamd_whiten_clips", _ptr(clips), M, T, L, _ptr(Wcm), _ptr(out), 0.0, None,
               _ptr(ws), ws.numel(), _stream(dev))
    return out


def whiten(X, quantization_unit=0.0):
    """ms3.whiten: covariance, whitening matrix, apply (p_whiten.cpp:45-206)."""
    C = whiten_covariance(X)
    W = whitening_matrix(C)
    return whiten_apply(X, W, quantization_unit)


# --------------------------------------------------------------------------
# artifact masking (a7)
# --------------------------------------------------------------------------
def mask_out_artifacts(X, threshold=6.0, interval_size=2000, return_use_it=False, out=None):
    """ms3.mask_out_artifacts (p_mask_out_artifacts.cpp:24-104)."""
    _check_ts(X)
    N, M = X.shape
    dev = X.device
    isz = int(interval_size)
    if threshold == 0 or isz == 0:
        raise ValueError("Problem with input parameters. Either threshold or interval_size is zero.")
    L = _capi.lib()
    ws = _workspace(L.msamd_mask_workspace_size(M, N, isz), dev)
    Y = out if out is not None else torch.empty((N, M), dtype=torch.float32, device=dev)
    use_it = torch.empty(N // isz + 1, dtype=torch.int32, device=dev)
    _capi.call("msamd_mask_out_artifacts", _ptr(X), M, N, float(threshold), isz, _ptr(Y),
               _ptr(use_it), _ptr(ws), ws.numel(), _stream(dev))
    return (Y, use_it) if return_use_it else Y


def mask_rows(X, threshold=6.0, interval_size=2000):
    """The decisions of mask_out_artifacts without the copy: returns
    (row_keep, use_it), row_keep an int32 tensor of ceil(N/32) bitmap words
    (bit t%32 of word t/32 set when row t is kept) for the row_keep argument
    of whiten_covariance and whiten_apply."""
    _check_ts(X)
    N, M = X.shape
    dev = X.device
    isz = int(interval_size)
    if threshold == 0 or isz == 0:
        raise ValueError("Problem with input parameters. Either threshold or interval_size is zero.")
    L = _capi.lib()
    ws = _workspace(L.msamd_mask_workspace_size(M, N, isz), dev)
    rk = torch.zeros(max((N + 31) // 32, 1), dtype=torch.int32, device=dev)
    use_it = torch.empty(N // isz + 1, dtype=torch.int32, device=dev)
    _capi.call("msamd_mask_rows", _ptr(X), M, N, float(threshold), isz, _ptr(rk), _ptr(use_it),
               _ptr(ws), ws.numel(), _stream(dev))
    return rk, use_it


# --------------------------------------------------------------------------
# Templates and the fit stage (p_compute_templates.cpp, globaltemplatecomputer.cpp,
# fitstagecomputer.cpp, fit_stage.cpp)
# --------------------------------------------------------------------------
def time_chunk_size(M, N, num_simultaneous, min_num_chunks):
    """get_time_chunk_infos chunk length (p_mountainsort3.cpp:507-533)."""
    return int(_capi.lib().msamd_time_chunk_size(int(M), int(N), int(num_simultaneous),
                                                  int(min_num_chunks)))


def _events_host(times, labels):
    import numpy as np
    t = np.ascontiguousarray(torch.as_tensor(times).cpu().numpy(), dtype=np.int64)
    lab = np.ascontiguousarray(torch.as_tensor(labels).cpu().numpy(), dtype=np.int32)
    if t.shape != lab.shape:
        raise ValueError("times and labels must have one length")
    return t, lab


def compute_templates(X, times, labels, K, clip_size):
    """ms3.compute_templates (p_compute_templates.cpp:24-118): the mean clip of
    every label 1..K over all channels, FP64 sums in event order.  X: (N, M)
    float32 on the device; times: 0-based sample indices; labels 1..K (others
    are skipped).  Returns (K, T, M) float32, the M x T x K .mda memory."""
    _check_ts(X)
    N, M = X.shape
    dev = X.device
    t, lab = _events_host(times, labels)
    K, T = int(K), int(clip_size)
    out = torch.zeros((max(K, 0), T, M), dtype=torch.float32, device=dev)
    if K == 0:
        return out
    L = _capi.lib()
    ws = _workspace(L.msamd_templates_workspace_size(len(t), K), dev)
    _capi.call("msamd_compute_templates", _ptr(X), M, N, t.ctypes.data_as(ctypes.c_void_p),
               lab.ctypes.data_as(ctypes.c_void_p), len(t), K, T, _ptr(out), _ptr(ws),
               ws.numel(), _stream(dev))
    return out


def fit_stage(X, times, labels, templates, chunk_size=None, num_threads=8, padding=1000,
              mask_threshold=0.1, return_passes=False):
    """The ms3alg fit stage (fitstagecomputer.cpp:55-104, fit_stage.cpp:27-275):
    X (N, M) float32 on the device, events with ascending times and labels
    1..K, templates (K, T, M) on the device.  chunk_size defaults to the
    reference plan get_time_chunk_infos(M, N, threads, threads).  Returns a
    numpy bool array: which events the fit stage keeps."""
    import numpy as np
    _check_ts(X)
    N, M = X.shape
    dev = X.device
    t, lab = _events_host(times, labels)
    K, T = int(templates.shape[0]), int(templates.shape[1])
    if templates.shape[2] != M or templates.device != dev:
        raise ValueError("templates must be (K, T, M) on the timeseries' device")
    tm = templates.to(torch.float32).contiguous()
    cs = int(chunk_size) if chunk_size else time_chunk_size(M, N, num_threads, num_threads)
    L = _capi.lib()
    pt = t.ctypes.data_as(ctypes.c_void_p)
    nb = L.msamd_fit_stage_workspace_size(M, N, pt, len(t), K, T, cs, int(padding))
    if nb < 0:
        raise _capi.MsamdError("msamd_fit_stage_workspace_size", 1,
                               "invalid events (times must be ascending) or shapes")
    ws = _workspace(nb, dev)
    keep = np.zeros(max(len(t), 1), np.uint8)
    passes = ctypes.c_int32(0)
    _capi.call("msamd_fit_stage", _ptr(X), M, N, pt, lab.ctypes.data_as(ctypes.c_void_p), len(t),
               _ptr(tm), K, T, float(mask_threshold), cs, int(padding),
               keep.ctypes.data_as(ctypes.c_void_p), ctypes.byref(passes), _ptr(ws), ws.numel(),
               _stream(dev))
    keep = keep[:len(t)].astype(bool)
    return (keep, passes.value) if return_passes else keep


# --------------------------------------------------------------------------
# Branch re-sorting: sort_clips (packages/mountainsortalg/sort_clips.cpp:29-125)
# --------------------------------------------------------------------------
def pca_subsets(data, indices, offsets, num_features=10, max_pca_samples=10000):
    """PCA features of index subsets of one (L, R) float32 device array (the
    R x L .mda memory): subset g is columns indices[offsets[g]:offsets[g+1]].
    Returns (features_flat, feature_offsets): subset g's (L_g, nf) features
    at features_flat[feature_offsets[g]:]."""
    import numpy as np
    if data.dim() != 2 or data.dtype != torch.float32 or not data.is_contiguous():
        raise ValueError("data must be a contiguous (L, R) float32 tensor")
    L, R = data.shape
    dev = data.device
    off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
    G = len(off) - 1
    idx = torch.as_tensor(indices, dtype=torch.int64).to(dev).contiguous()
    nf = int(num_features)
    lib = _capi.lib()
    po = off.ctypes.data_as(ctypes.c_void_p)
    nb = lib.msamd_pca_subsets_workspace_size(G, po, R, nf, int(max_pca_samples))
    if nb < 0:
        raise _capi.MsamdError("msamd_pca_subsets_workspace_size", 1, "unsupported subsets")
    ws = _workspace(nb, dev)
    foff = off * nf
    feats = torch.empty(max(int(foff[-1]), 1), dtype=torch.float32, device=dev)
    _capi.call("msamd_pca_subsets", _ptr(data), R, L, _ptr(idx), po, G, nf,
               int(max_pca_samples), _ptr(feats), _ptr(ws), ws.numel(), _stream(dev))
    return feats, foff


def sort_clips(clips, cluster, num_features=10, max_pca_samples=10000, max_depth=64):
    """sort_clips with the branch method (sort_clips.cpp:29-125), the
    clustering supplied by the caller (isosplit5 in the reference):
    cluster(features) takes an (L0, nf) float32 numpy array and returns L0
    integer labels (1..K0; 0 = unassigned).  clips: (L, T, M) float32 on the
    device (the M x T x L reduced clips).  Each level of the recursion runs
    the PCA of all its subsets in one batched launch; the labels are composed
    in the reference's depth-first order (the sub-labels of cluster k offset
    by the clusters found under clusters 1..k-1).  Returns (L,) int labels."""
    import numpy as np
    L = int(clips.shape[0])
    data = clips.reshape(L, -1).contiguous()
    root = {"idx": np.arange(L, dtype=np.int64), "children": None, "labels": None}
    level = [root]
    depth = 0
    while level:
        depth += 1
        if depth > max_depth:
            raise RuntimeError(f"sort_clips: recursion deeper than {max_depth}")
        off = np.concatenate([[0], np.cumsum([len(n["idx"]) for n in level])]).astype(np.int64)
        feats, foff = pca_subsets(data, np.concatenate([n["idx"] for n in level]), off,
                                  num_features, max_pca_samples)
        feats = feats.cpu().numpy()
        nxt = []
        for g, node in enumerate(level):
            Lg = len(node["idx"])
            F = feats[foff[g]:foff[g + 1]].reshape(Lg, num_features)
            lab = np.asarray(cluster(F), dtype=np.int64).reshape(Lg)
            node["labels"] = lab
            K0 = int(lab.max()) if Lg else 0
            if K0 > 1:
                node["children"] = []
                for k in range(1, K0 + 1):
                    sel = np.nonzero(lab == k)[0]
                    child = {"idx": node["idx"][sel], "pos": sel, "children": None,
                             "labels": None} if len(sel) else None
                    node["children"].append(child)
                    if child is not None:
                        nxt.append(child)
        level = nxt

    def resolve(node):
        lab = node["labels"]
        if node["children"] is None:
            return lab
        out = np.zeros(len(lab), np.int64)
        koff = 0
        for child in node["children"]:
            if child is None:
                continue
            sub = resolve(child)
            out[child["pos"]] = koff + sub
            koff += int(sub.max()) if len(sub) else 0
        return out

    return resolve(root)
