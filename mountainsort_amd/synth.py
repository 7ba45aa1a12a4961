"""Seeded synthetic extracellular recordings for tests and benchmarks.

Waveform model after packages/pyms/synthesis/p_synthesize_timeseries.py:72-89
(templates added at spike times over noise): every unit has a biphasic
T-sample waveform (trough at sample 20 with Gaussian width U(1.5, 3.5) and
depth U(amp_lo, amp_hi), a positive lobe of 0.2-0.5x the depth U(6, 12)
samples later with width U(3, 7)), a spatial gain per channel and Poisson
firing.  Noise is white Gaussian or pink+white (1/f shaped in blocks), total
sigma ``noise_sigma``.  Optional common-mode artifact bursts and int16
quantization.  Everything is driven by numpy's PCG64 with the given seed, so a
recording is reproducible from (config, seed).

Returned tensor: (N, M) float32 or int16, i.e. the .mda M x N memory order.
"""
from dataclasses import dataclass, field

import numpy as np
import torch


@dataclass
class SynthConfig:
    M: int = 4
    N: int = 1_800_000
    samplerate: float = 30000.0
    num_units: int = 5
    amp_lo: float = 50.0
    amp_hi: float = 200.0
    rate_lo: float = 5.0            # Hz; equal lo/hi -> fixed rate
    rate_hi: float = 5.0
    log_uniform_rate: bool = False
    noise_sigma: float = 10.0
    pink: bool = False
    geometry: str = "none"          # "none" (random per-channel gains) | "probe"
    columns: int = 2
    pitch: float = 20.0
    footprint_sigma: float = 40.0   # um, Gaussian spatial footprint
    artifact_fraction: float = 0.0
    artifact_amp: float = 2000.0
    int16_scale: float = 0.0        # uV per bit; > 0 -> int16 counts
    clip_size: int = 50
    seed: int = 0
    extra: dict = field(default_factory=dict)


def config0(N=1_800_000, seed=0):
    """4-channel, 60 s at 30 kHz, 5 units, white noise sigma 10."""
    return SynthConfig(M=4, N=N, num_units=5, amp_lo=50, amp_hi=200, rate_lo=5, rate_hi=5,
                       seed=seed)


def config1(N=18_000_000, seed=1):
    """64-channel 2-column probe (20 um pitch), 40 units, pink+white noise."""
    return SynthConfig(M=64, N=N, num_units=40, amp_lo=40, amp_hi=300, rate_lo=1, rate_hi=20,
                       log_uniform_rate=True, pink=True, geometry="probe", columns=2, seed=seed)


def config2(N=9_000_000, seed=2):
    """384-channel 4-column probe, 300 units, artifacts on 1% of samples, int16 0.195 uV/bit."""
    return SynthConfig(M=384, N=N, num_units=300, amp_lo=40, amp_hi=300, rate_lo=1, rate_hi=20,
                       log_uniform_rate=True, pink=True, geometry="probe", columns=4,
                       artifact_fraction=0.01, int16_scale=0.195, seed=seed)


def probe_geometry(M, columns, pitch):
    """(M, 2) electrode positions: `columns` columns x M/columns rows at `pitch` um."""
    idx = np.arange(M)
    x = (idx % columns) * pitch
    y = (idx // columns) * pitch
    return np.stack([x, y], axis=1).astype(np.float64)


def _waveforms(rng, K, T, amp_lo, amp_hi):
    t = np.arange(T, dtype=np.float64)
    W = np.zeros((K, T))
    for k in range(K):
        depth = rng.uniform(amp_lo, amp_hi)
        w1 = rng.uniform(1.5, 3.5)
        lobe = rng.uniform(0.2, 0.5) * depth
        dt = rng.uniform(6, 12)
        w2 = rng.uniform(3, 7)
        W[k] = -depth * np.exp(-0.5 * ((t - 20) / w1) ** 2) + lobe * np.exp(-0.5 * ((t - 20 - dt) / w2) ** 2)
    return W


def _gains(rng, cfg):
    if cfg.geometry == "probe":
        geom = probe_geometry(cfg.M, cfg.columns, cfg.pitch)
        span_y = geom[:, 1].max()
        pos = np.stack([rng.uniform(0, geom[:, 0].max() + 1e-9, cfg.num_units),
                        rng.uniform(0, span_y, cfg.num_units)], axis=1)
        d2 = ((pos[:, None, :] - geom[None, :, :]) ** 2).sum(-1)
        G = np.exp(-d2 / (2 * cfg.footprint_sigma ** 2))
        G[G < 1e-3] = 0.0
        return G
    return rng.uniform(0.2, 1.0, size=(cfg.num_units, cfg.M))


def _spike_times(rng, cfg):
    dur = cfg.N / cfg.samplerate
    times = []
    for _ in range(cfg.num_units):
        if cfg.log_uniform_rate:
            rate = float(np.exp(rng.uniform(np.log(cfg.rate_lo), np.log(cfg.rate_hi))))
        else:
            rate = float(rng.uniform(cfg.rate_lo, cfg.rate_hi))
        n = rng.poisson(rate * dur)
        st = np.sort(rng.integers(0, max(cfg.N - cfg.clip_size, 1), size=n))
        times.append(st.astype(np.int64))
    return times


def generate(cfg, device="cpu", block=1 << 20, return_truth=False):
    """Generate the recording described by cfg on `device`."""
    rng = np.random.Generator(np.random.PCG64(cfg.seed))
    dev = torch.device(device)
    T = cfg.clip_size
    wave = _waveforms(rng, cfg.num_units, T, cfg.amp_lo, cfg.amp_hi)
    gains = _gains(rng, cfg)
    spikes = _spike_times(rng, cfg)
    noise_seed = int(rng.integers(0, 2**62))
    X = torch.empty((cfg.N, cfg.M), dtype=torch.float32, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(noise_seed)
    # noise, block by block
    for a in range(0, cfg.N, block):
        b = min(a + block, cfg.N)
        n = b - a
        white = torch.randn((n, cfg.M), generator=gen, device=dev, dtype=torch.float32)
        if cfg.pink:
            nf = 1 << int(np.ceil(np.log2(max(n, 2))))
            src = torch.randn((nf, cfg.M), generator=gen, device=dev, dtype=torch.float32)
            F = torch.fft.rfft(src, dim=0)
            f = torch.arange(F.shape[0], device=dev, dtype=torch.float32)
            f[0] = 1.0
            F = F / torch.sqrt(f)[:, None]
            F[0] = 0
            pink = torch.fft.irfft(F, n=nf, dim=0)[:n]
            pink = pink / pink.std().clamp_min(1e-12)
            noise = (white + pink) * (cfg.noise_sigma / np.sqrt(2.0))
        else:
            noise = white * cfg.noise_sigma
        X[a:b] = noise
    # spikes: per unit template (T, M), added at each spike time
    for k in range(cfg.num_units):
        st = spikes[k]
        if len(st) == 0:
            continue
        tmpl = torch.tensor(np.outer(wave[k], gains[k]), dtype=torch.float32, device=dev)
        idx = torch.tensor(st, device=dev)
        for s in range(T):
            X.index_add_(0, idx + s, tmpl[s].expand(len(st), cfg.M))
    # artifacts: common-mode bursts of random sign covering artifact_fraction
    if cfg.artifact_fraction > 0:
        total = int(cfg.artifact_fraction * cfg.N)
        covered = 0
        while covered < total:
            dur = int(rng.uniform(0.005, 0.050) * cfg.samplerate)
            t0 = int(rng.integers(0, max(cfg.N - dur, 1)))
            sign = 1.0 if rng.random() < 0.5 else -1.0
            X[t0:t0 + dur] += sign * cfg.artifact_amp
            covered += dur
    if cfg.int16_scale > 0:
        Xi = torch.empty((cfg.N, cfg.M), dtype=torch.int16, device=dev)
        for a in range(0, cfg.N, block):
            b = min(a + block, cfg.N)
            Xi[a:b] = torch.round(X[a:b] / cfg.int16_scale).clamp_(-32768, 32767).to(torch.int16)
        del X
        X = Xi
    if return_truth:
        return X, {"spike_times": spikes, "waveforms": wave, "gains": gains}
    return X


# ---------------------------------------------------------------------------
# Clip batches for per-neighborhood PCA
# ---------------------------------------------------------------------------
def pca_batch_shapes(G=384, seed=3, w_lo=5, w_hi=12, l_lo=100, l_hi=200_000):
    """Neighborhood shapes of the batched PCA workload: widths uniform in
    [w_lo, w_hi] channels and event counts log-uniform in [l_lo, l_hi].
    Returns (num_events int64 (G,), widths int32 (G,))."""
    rng = np.random.default_rng(seed)
    W = rng.integers(w_lo, w_hi + 1, size=G).astype(np.int32)
    L = np.round(np.exp(rng.uniform(np.log(l_lo), np.log(l_hi), size=G))).astype(np.int64)
    return L, W


def pca_batch_clips(num_events, widths, T=50, seed=3, units=4, noise_sigma=1.0, device="cpu"):
    """Packed clips for the batched PCA (the extract_clips_batched layout:
    neighborhood g holds L_g x T x W_g floats, element (i, t, w) at
    i*T*W + t*W + w).  Each neighborhood has `units` biphasic waveforms
    (the recording model above, amplitudes 3-12 noise sigmas) with spatial
    gains U(0.1, 1); a clip is one of them at amplitude U(0.7, 1.3) plus
    white noise.  Templates and labels come from numpy (seeded); the bulk
    noise is drawn on `device` with torch's generator seeded from `seed`.
    Returns (clips_flat float32, clip_offsets int64 (G+1,))."""
    rng = np.random.default_rng(seed + 1000)
    dev = torch.device(device)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    L = np.asarray(num_events, dtype=np.int64)
    W = np.asarray(widths, dtype=np.int64)
    sizes = L * T * W
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    out = torch.empty(max(int(off[-1]), 1), dtype=torch.float32, device=dev)
    for g in range(len(L)):
        if L[g] == 0:
            continue
        wave = _waveforms(rng, units, T, 3.0 * noise_sigma, 12.0 * noise_sigma)  # (K, T)
        gain = rng.uniform(0.1, 1.0, size=(units, W[g]))
        temps = torch.from_numpy((wave[:, :, None] * gain[:, None, :]).astype(np.float32)).to(dev)
        lab = torch.from_numpy(rng.integers(0, units, L[g])).to(dev)
        amp = torch.from_numpy(rng.uniform(0.7, 1.3, L[g]).astype(np.float32)).to(dev)
        seg = out[off[g]:off[g + 1]].view(int(L[g]), T, int(W[g]))
        torch.randn(seg.shape, generator=gen, device=dev, out=seg)
        seg.mul_(noise_sigma).add_(temps[lab] * amp[:, None, None])
    return out, off
