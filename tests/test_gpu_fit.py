"""GPU parity: templates and the fit stage against the oracle's restatement
of p_compute_templates / GlobalTemplateComputer / FitStageComputer /
fit_stage.  Templates must be bit-identical (FP64 sums in the reference's
event order); the set of events the fit stage keeps must be identical."""
import numpy as np
import pytest
import torch

from oracle import core as oc
from tests.conftest import say

pytestmark = pytest.mark.gpu


def _sorted_events(M, N, seed, units, extra_false=40):
    """A recording with `units` units (synth model, white noise), events at
    the true spike centers labeled by unit, plus unlabeled and false events,
    duplicates and an event at chunk-local time 23 of a chunk."""
    from mountainsort_amd import synth
    cfg = synth.SynthConfig(M=M, N=N, num_units=units, amp_lo=40, amp_hi=150, rate_lo=3,
                            rate_hi=15, seed=seed)
    X, truth = synth.generate(cfg, device="cpu", return_truth=True)
    X = (X.numpy() / 10.0).astype(np.float32)    # unit noise, so the 0.1 mask threshold bites
    rng = np.random.default_rng(seed)
    times, labels = [], []
    for k, st in enumerate(truth["spike_times"]):
        times.extend((st + 20).tolist())          # trough at clip sample 24
        labels.extend([k + 1] * len(st))
    times = np.array(times, np.int64)
    labels = np.array(labels, np.int32)
    ft = rng.integers(30, N - 30, extra_false)
    fl = rng.integers(0, units + 1, extra_false).astype(np.int32)   # label 0: skipped
    dup = rng.choice(len(times), 10, replace=False)
    times = np.concatenate([times, ft, times[dup]])
    labels = np.concatenate([labels, fl, labels[dup]])
    o = np.argsort(times, kind="stable")
    return X, times[o], labels[o]


@pytest.mark.parametrize("M,N,units", [(16, 600_000, 12), (64, 300_000, 30), (5, 90_000, 3)])
def test_templates_and_fit_stage_match_the_oracle(gpu, M, N, units):
    from mountainsort_amd import frontend as fe
    X, times, labels = _sorted_events(M, N, seed=M + units, units=units)
    T, K = 50, units
    cs = oc.time_chunk_infos(M, N, 8, 8)
    if len(cs) > 1:   # an event at chunk-local time 23 of the second chunk
        t23 = cs[1][0] - cs[1][2] + 23
        i = np.searchsorted(times, t23)
        times = np.insert(times, i, t23)
        labels = np.insert(labels, i, 1)
    say(f"templates + fit stage M={M} N={N} units={units}: {len(times)} events, "
        f"{len(cs)} chunks")
    Xg = torch.from_numpy(X).to(gpu)
    # ms3.compute_templates over every event
    tm_o, cnt = oc.compute_templates(X, times, labels, K, T)
    tm_g = fe.compute_templates(Xg, times, labels, K, T).cpu().numpy()
    ndiff = int(np.sum(tm_g != tm_o))
    say(f"  compute_templates: {ndiff} of {tm_o.size} entries differ")
    assert np.array_equal(tm_g, tm_o)
    # ms3alg: global templates over the chunk rule, then the fit stage
    sel = oc.global_template_events(times, N, M)
    gt_o, _ = oc.global_templates(X, times, labels, K, T)
    gt_g = fe.compute_templates(Xg, times[sel], labels[sel], K, T)
    assert np.array_equal(gt_g.cpu().numpy(), gt_o)
    used_o = oc.fit_stage_computer(X, times, labels, gt_o, 8)
    keep, passes = fe.fit_stage(Xg, times, labels, gt_g, num_threads=8, return_passes=True)
    used_g = np.nonzero(keep)[0]
    say(f"  fit stage: {len(used_g)} of {len(times)} kept (oracle {len(used_o)}), "
        f"at most {passes} passes per chunk")
    assert np.array_equal(used_g, used_o)
    assert passes >= 2
    assert 0 < len(used_g) < len(times)


def test_fit_stage_single_chunk_against_fit_stage(gpu):
    """One chunk spanning the recording with no padding: the kept set equals
    fit_stage's on the same data, and the input timeseries is not modified."""
    from mountainsort_amd import frontend as fe
    X, times, labels = _sorted_events(8, 100_000, seed=5, units=5)
    tm_o, _ = oc.compute_templates(X, times, labels, 5, 50)
    keep_o, p_o = oc.fit_stage(X.copy(), times.astype(np.float64), labels, tm_o)
    Xg = torch.from_numpy(X).to(gpu)
    keep_g, p_g = fe.fit_stage(Xg, times, labels, torch.from_numpy(tm_o).to(gpu),
                               chunk_size=100_000, padding=0, return_passes=True)
    say(f"fit stage, one chunk: {keep_g.sum()} kept (oracle {keep_o.sum()}), passes {p_g} ({p_o})")
    assert np.array_equal(keep_g, keep_o)
    assert p_g == p_o
    assert torch.equal(Xg.cpu(), torch.from_numpy(X))


def test_fit_stage_rejects_unsorted_times(gpu):
    from mountainsort_amd import _capi
    from mountainsort_amd import frontend as fe
    X = torch.zeros((1000, 2), device=gpu)
    tm = torch.zeros((1, 50, 2), device=gpu)
    with pytest.raises(_capi.MsamdError):
        fe.fit_stage(X, np.array([300, 100]), np.array([1, 1]), tm, chunk_size=1000)
