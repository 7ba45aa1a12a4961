"""GPU parity: batched two-stage PCA features against the CPU oracle (the
reference's float-form PCA) and the FP64 Gram restatement."""
import numpy as np
import pytest
import torch

from oracle import core as oc
from tests.conftest import assert_pca_parity, say

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
        e1, e2 = assert_pca_parity(r_g, r_o, f_g, f_o, f"g={g} L={L} W={W}", clips=c)
        worst1, worst2 = max(worst1, e1), max(worst2, e2)
    say(f"  worst stage1 {worst1:.2e}, stage2 {worst2:.2e}")


def test_pca_plan_graph_and_per_neighborhood_match(gpu):
    """The reusable plan, its HIP-graph replay, one plan per neighborhood and
    the one-call entry point give bit-identical features (every neighborhood's
    work items are the same in all four), and the graph stays correct over
    replays with other device work in between."""
    from mountainsort_amd import frontend as fe
    from mountainsort_amd import synth
    L, W = synth.pca_batch_shapes(G=24, seed=11, l_hi=30_000)
    L[3] = 0
    say(f"pca plan: G=24, events {int(L.sum())}, widths {W.min()}-{W.max()}")
    clips, co = synth.pca_batch_clips(L, W, seed=11, device=gpu)
    eo = np.concatenate([[0], np.cumsum(L)])
    no = np.concatenate([[0], np.cumsum(W)])
    red1, ro, feat1, fo = fe.pca_features(clips, co, eo, no, 50)
    red1, feat1 = red1.clone(), feat1.clone()
    plan = fe.PcaPlan(L, W, 50, device=gpu)
    assert plan.tiles == (2560, 64, 64)
    red2, ro2, feat2, fo2 = plan.run(clips)
    assert np.array_equal(ro, ro2) and np.array_equal(fo, fo2)
    assert torch.equal(red1, red2) and torch.equal(feat1, feat2)
    for rep in range(3):
        plan.reduced.fill_(float("nan"))
        plan.features.fill_(float("nan"))
        torch.cuda.synchronize()
        _, _, feat3, _ = plan.run(clips, graph=True)
        torch.cuda.synchronize()
        say(f"  graph replay {rep}")
        assert torch.equal(plan.reduced, red1) and torch.equal(feat3, feat1)
        junk = torch.randn(1 << 20, device=gpu).sum()  # eager work between replays
        assert torch.isfinite(junk)
    plan.close()
    for g in range(len(L)):
        p1 = fe.PcaPlan(L[g:g + 1], W[g:g + 1], 50, device=gpu)
        seg = clips[co[g]:co[g + 1]].contiguous() if L[g] else torch.zeros(1, device=gpu)
        r, _, f, _ = p1.run(seg)
        assert torch.equal(r[:ro[g + 1] - ro[g]], red1[ro[g]:ro[g + 1]])
        assert torch.equal(f[:fo[g + 1] - fo[g]], feat1[fo[g]:fo[g + 1]])
        p1.close()


def test_pca_autotune_tiles_keep_parity(gpu):
    """The autotuner returns tiles from its candidate sets, and a plan with
    them still matches the CPU oracle on every component and feature."""
    from mountainsort_amd import frontend as fe
    rng = np.random.default_rng(5)
    shapes = [(20000, 7), (700, 12), (3000, 5)]
    say(f"pca autotune on (L, W) = {shapes}")
    host = [_clips(rng, L, W) for L, W in shapes]
    flat = torch.from_numpy(np.concatenate([c.ravel() for c in host])).to(gpu)
    L = np.array([s[0] for s in shapes])
    W = np.array([s[1] for s in shapes])
    tiles, ms = fe.pca_autotune(flat, L, W, 50)
    say(f"  tiles {tiles}, {ms:.3f} ms")
    assert tiles[0] in (1280, 2560, 5120, 10240)
    assert tiles[1] in (32, 64, 128, 256) and tiles[2] in (16, 32, 64, 128)
    assert ms > 0
    plan = fe.PcaPlan(L, W, 50, tiles=tiles, device=gpu)
    red, ro, feat, fo = plan.run(flat)
    red = red.cpu().numpy()
    feat = feat.cpu().numpy()
    for g, ((Lg, Wg), c) in enumerate(zip(shapes, host)):
        r_o = oc.dimension_reduce_clips(c, 10, 10000)
        f_o = oc.stage2_features(r_o, 10, 10000)
        r_g = red[ro[g]:ro[g + 1]].reshape(Lg, 10, Wg)
        f_g = feat[fo[g]:fo[g + 1]].reshape(Lg, 10)
        assert_pca_parity(r_g, r_o, f_g, f_o, f"g={g} tiles {tiles}", clips=c)
    plan.close()


@pytest.mark.parametrize("shapes", [
    [(12000, 64), (3000, 40), (700, 17)],
    [(5000, 100), (20000, 128), (40, 64)],
])
def test_pca_wide_neighborhoods(gpu, shapes):
    """Neighborhoods wider than one 16-channel work item (stage-1 channel
    groups) and stage-2 dimensions above 256 (band pairs of the stage-2 Gram,
    its FP64 copy in the workspace): the ms3alg default adjacency_radius = -1
    puts all 64 channels of a 64-channel probe in every neighborhood."""
    from mountainsort_amd import frontend as fe
    rng = np.random.default_rng(sum(L * W for L, W in shapes))
    say(f"pca wide neighborhoods (L, W) = {shapes}")
    clips = [_clips(rng, L, W) for L, W in shapes]
    flat = torch.from_numpy(np.concatenate([c.ravel() for c in clips])).to(gpu)
    eo = np.concatenate([[0], np.cumsum([L for L, _ in shapes])])
    no = np.concatenate([[0], np.cumsum([W for _, W in shapes])])
    co = np.concatenate([[0], np.cumsum([L * W * 50 for L, W in shapes])])
    red, ro, feat, fo = fe.pca_features(flat, co, eo, no, 50)
    red, feat = red.cpu().numpy(), feat.cpu().numpy()
    for g, ((L, W), c) in enumerate(zip(shapes, clips)):
        r_o = oc.dimension_reduce_clips(c, 10, 10000)
        f_o = oc.stage2_features(r_o, 10, 10000)
        assert_pca_parity(red[ro[g]:ro[g + 1]].reshape(L, 10, W), r_o,
                          feat[fo[g]:fo[g + 1]].reshape(L, 10), f_o, f"g={g} L={L} W={W}",
                          clips=c)


def _reduced_clusters(rng, sizes, W, nf1=10, sep=6.0):
    """Reduced clips (L, nf1, W) of well-separated clusters of the given sizes."""
    parts = [rng.normal(size=(nf1, W)) * sep + rng.normal(size=(n, nf1, W)) for n in sizes]
    X = np.concatenate(parts).astype(np.float32)
    truth = np.repeat(np.arange(len(sizes)), sizes)
    o = rng.permutation(len(truth))
    return np.ascontiguousarray(X[o]), truth[o]


@pytest.mark.parametrize("W", [4, 16, 64])
def test_pca_subsets_match_oracle(gpu, W):
    """Every subset's features against pca_subsampled on that subset (the
    reference's float form) and the FP64 Gram form, including empty, one-event
    and subsampled (L0 > 10000) subsets."""
    from mountainsort_amd import frontend as fe
    from tests.conftest import PCA_STAGE2_BOUND, _rel
    rng = np.random.default_rng(W)
    L = 24000
    X, _ = _reduced_clusters(rng, [6000, 9000, 5000, 4000], W)
    sizes = [0, 1, 7, 300, 12001, 2500]
    idx = [np.sort(rng.choice(L, n, replace=False)) if n else np.zeros(0, np.int64)
           for n in sizes]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    say(f"pca subsets W={W} R={10 * W} sizes={sizes}")
    feats, foff = fe.pca_subsets(torch.from_numpy(X.reshape(L, -1)).to(gpu),
                                 np.concatenate(idx), off)
    feats = feats.cpu().numpy()
    for g, (n, ix) in enumerate(zip(sizes, idx)):
        if n == 0:
            assert foff[g + 1] == foff[g]
            continue
        sub = X[ix].reshape(n, -1)
        f_g = feats[foff[g]:foff[g + 1]].reshape(n, 10)
        _, f_o, _ = oc.pca_subsampled(sub, 10, 10000)
        _, f_64, _ = oc.pca_gram_fp64(sub, 10, 10000)
        # a subset of n < 10 events spans n dimensions: the later components
        # come from the power method run on the rounding residue of the
        # deflation and are arbitrary, in the reference as here
        assert np.all(np.isfinite(f_g))
        k = min(n, 10)
        f_g, f_o, f_64 = f_g[:, :k], f_o[:, :k], f_64[:, :k]
        e, g_o, o = _rel(f_g, f_64), _rel(f_g, f_o), _rel(f_o, f_64)
        say(f"  subset {g} L0={n}: vs FP64 {e:.2e}, vs oracle {g_o:.2e} (oracle vs FP64 {o:.2e})")
        assert e < PCA_STAGE2_BOUND
        assert g_o < PCA_STAGE2_BOUND + o


@pytest.mark.parametrize("W,sizes", [(8, [900, 400, 1200, 300, 700, 500]),
                                     (32, [5000, 6000, 3000, 800, 2500, 4000, 1500, 900])])
def test_sort_clips_branch_method_matches_oracle(gpu, W, sizes):
    """The branch-method recursion: labels identical to the oracle's
    depth-first restatement with the same caller clustering, and the
    separated clusters recovered."""
    from mountainsort_amd import frontend as fe
    from tests.test_oracle import gap_split
    rng = np.random.default_rng(7 + W)
    X, truth = _reduced_clusters(rng, sizes, W)
    say(f"sort_clips W={W} L={len(truth)} clusters={len(sizes)}")
    calls = []

    def cluster(F):
        calls.append(len(F))
        return gap_split(F)

    lab_g = fe.sort_clips(torch.from_numpy(X).to(gpu), cluster)
    lab_o = oc.sort_clips(X, gap_split)
    say(f"  {len(calls)} clusterings, {lab_g.max()} clusters (oracle {lab_o.max()})")
    assert np.array_equal(lab_g, lab_o)
    pairs = {(int(a), int(b)) for a, b in zip(lab_g, truth)}
    assert len(pairs) == len(set(lab_g.tolist())) == len(sizes)


def _gather_case(rng, M=64, N=120_000, T=50):
    """Neighborhoods of several widths (central channel first) with sorted
    event times that include clips running past both ends of the recording,
    and one neighborhood above max_pca_samples (subsampled Gram columns)."""
    shapes = [(300, 1), (12345, 5), (900, 12), (40, 13), (2500, 17), (0, 3), (700, 40)]
    nb, times = [], []
    edge = np.array([0, 1, 5, 23, 24, 25, N - 26, N - 25, N - 2, N - 1])
    for L, W in shapes:
        c = int(rng.integers(0, M))
        others = rng.choice(np.setdiff1d(np.arange(M), [c]), W - 1, replace=False)
        nb.append(np.concatenate([[c], np.sort(others)]).astype(np.int32))
        t = rng.integers(0, N, max(L - len(edge), 0))
        t = np.sort(np.concatenate([t, edge[:min(L, len(edge))]]))
        times.append(t.astype(np.int64))
    return shapes, nb, times


@pytest.mark.parametrize("tiles", [None, (1280, 32, 16), (10240, 256, 128)])
def test_pca_fused_gather_matches_clip_path(gpu, tiles):
    """The fused form (stage-1 kernels reading clips from the timeseries)
    gives features bit-identical to the clip path with the same tiles, edge
    clips (zero-filled outside the recording) and subsampled columns
    included."""
    from mountainsort_amd import frontend as fe
    rng = np.random.default_rng(17)
    M, N = 64, 120_000
    X = torch.from_numpy(rng.normal(size=(N, M)).astype(np.float32) * 5).to(gpu)
    shapes, nb, times = _gather_case(rng, M, N)
    say(f"pca fused gather: (L, W) = {shapes}, tiles {tiles}")
    eo = np.concatenate([[0], np.cumsum([len(t) for t in times])])
    no = np.concatenate([[0], np.cumsum([len(c) for c in nb])])
    t_d = torch.from_numpy(np.concatenate(times)).to(gpu)
    nc = np.concatenate(nb)
    clips, co = fe.extract_clips_batched(X, t_d, eo, nc, no, 50)
    L = eo[1:] - eo[:-1]
    W = no[1:] - no[:-1]
    plan = fe.PcaPlan(L, W, 50, tiles=tiles, device=gpu)
    red1, ro, feat1, fo = plan.run(clips)
    red1, feat1 = red1.clone(), feat1.clone()
    plan.reduced.fill_(float("nan"))
    plan.features.fill_(float("nan"))
    red2, _, feat2, _ = plan.run_timeseries(X, t_d, nc, no)
    assert torch.equal(red1, red2)
    assert torch.equal(feat1, feat2)
    plan.close()
    # the one-call form and the oracle on the neighborhood with edge clips
    red3, _, feat3, _ = fe.pca_features_from_timeseries(X, t_d, eo, nc, no, 50, tiles=tiles)
    assert torch.equal(red1, red3) and torch.equal(feat1, feat3)
    g = 2
    c = clips[co[g]:co[g + 1]].cpu().numpy().reshape(L[g], 50, W[g])
    r_o = oc.dimension_reduce_clips(c, 10, 10000)
    f_o = oc.stage2_features(r_o, 10, 10000)
    assert_pca_parity(red1[ro[g]:ro[g + 1]].cpu().numpy().reshape(L[g], 10, W[g]), r_o,
                      feat1[fo[g]:fo[g + 1]].cpu().numpy().reshape(L[g], 10), f_o,
                      f"fused g={g}", clips=c)


def test_pca_autotune_timeseries(gpu):
    """The fused form's autotuner returns tiles from the candidate sets, and
    the fused features at those tiles equal the clip path's."""
    from mountainsort_amd import frontend as fe
    rng = np.random.default_rng(23)
    M, N = 64, 120_000
    X = torch.from_numpy(rng.normal(size=(N, M)).astype(np.float32)).to(gpu)
    shapes, nb, times = _gather_case(rng, M, N)
    eo = np.concatenate([[0], np.cumsum([len(t) for t in times])])
    no = np.concatenate([[0], np.cumsum([len(c) for c in nb])])
    t_d = torch.from_numpy(np.concatenate(times)).to(gpu)
    nc = np.concatenate(nb)
    say("pca autotune (fused gather)")
    tiles, ms = fe.pca_autotune_timeseries(X, t_d, eo, nc, no, 50)
    say(f"  tiles {tiles}, {ms:.3f} ms")
    assert tiles[0] in (1280, 2560, 5120, 10240)
    assert tiles[1] in (32, 64, 128, 256) and tiles[2] in (16, 32, 64, 128)
    clips, co = fe.extract_clips_batched(X, t_d, eo, nc, no, 50)
    r1, _, f1, _ = fe.pca_features(clips, co, eo, no, 50, tiles=tiles)
    r2, _, f2, _ = fe.pca_features_from_timeseries(X, t_d, eo, nc, no, 50, tiles=tiles)
    assert torch.equal(r1, r2) and torch.equal(f1, f2)
