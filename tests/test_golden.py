"""Golden fixtures (tests/golden/front_end_small.npz, written by
tests/golden/make_golden.py): the oracle must keep reproducing them (CPU), and
the GPU path must match them to the same bars as the parity tests (GPU)."""
import os

import numpy as np
import pytest

from tests.golden import make_golden as mg

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "front_end_small.npz")


@pytest.fixture(scope="module")
def golden():
    with np.load(FIX, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.fixture(scope="module")
def inp():
    return mg.inputs()


def test_oracle_reproduces_fixtures(golden, inp):
    out = mg.expected(inp)
    assert set(out) == set(golden)
    for k, v in golden.items():
        if np.issubdtype(v.dtype, np.integer):
            assert np.array_equal(out[k], v), k
        elif k == "bandpass_val":
            # an FP64 DFT pair: its last bits depend on the FFT library's
            # factorization order, not on the filter
            rms = np.sqrt(np.mean(v ** 2))
            assert np.max(np.abs(out[k] - v)) < 1e-12 * rms, k
        else:
            np.testing.assert_allclose(out[k], v, rtol=1e-12, atol=0, err_msg=k)


@pytest.mark.gpu
def test_gpu_matches_fixtures(gpu, golden, inp):
    import torch

    from mountainsort_amd import frontend as fe
    from tests.conftest import assert_pca_parity, say

    say("golden: bandpass")
    bp = fe.bandpass_filter(torch.from_numpy(inp["raw"]).to(gpu), 30000, 300, 6000,
                            num_threads=94).cpu().numpy().ravel()
    ref = golden["bandpass_val"]
    err = np.max(np.abs(bp[golden["bandpass_idx"]] - ref)) / np.sqrt(np.mean(ref ** 2))
    assert err < 1e-4

    say("golden: mask, whitening")
    Y, use_it = fe.mask_out_artifacts(torch.from_numpy(inp["rec"]).to(gpu), 6.0, 2000,
                                      return_use_it=True)
    Yn = Y.cpu().numpy().astype(np.float64)
    assert np.array_equal(use_it.cpu().numpy()[:len(golden["mask_use_it"])],
                          golden["mask_use_it"])
    assert np.array_equal(np.array([Yn.sum(), np.abs(Yn).sum()]), golden["mask_checksum"])
    C = fe.whiten_covariance(Y).cpu().numpy()
    Cg = golden["whiten_cov"]
    assert np.max(np.abs(C - Cg)) / np.max(np.abs(np.diag(Cg))) < 2e-5
    W = fe.whitening_matrix(torch.from_numpy(Cg).to(gpu)).cpu().numpy()
    assert np.max(np.abs(W - golden["whiten_W"])) / np.max(np.abs(golden["whiten_W"])) < 1e-10
    Yw = fe.whiten_apply(Y, torch.from_numpy(golden["whiten_W"]).to(gpu)).cpu().numpy()
    q = np.rint(Yw[::997].astype(np.float64) / 1e-4) - np.rint(
        golden["whiten_rows"].astype(np.float64) / 1e-4)
    assert np.max(np.abs(q)) <= 1

    say("golden: detection, clips, PCA")
    off, times = fe.detect_events(torch.from_numpy(inp["white"]).to(gpu), 3.0, 10, -1,
                                  num_threads=8, chunk_size=20_000)
    assert np.array_equal(np.diff(off.cpu().numpy()), golden["detect_counts"])
    assert np.array_equal(times.cpu().numpy(), golden["detect_times"])
    n0 = int(golden["detect_counts"][0])
    clips = fe.extract_clips(torch.from_numpy(inp["white"]).to(gpu),
                             times[:n0] - (50 + 1) // 2 + 1, [0, 2, 4], 50).cpu().numpy()
    assert np.array_equal(clips, golden["clips_ch0"])
    c = inp["clips"]
    L, T, Wd = c.shape
    red, ro, feat, fo = fe.pca_features(torch.from_numpy(c.ravel()).to(gpu), [0, c.size],
                                        [0, L], [0, Wd], T)
    red = red.cpu().numpy()[:ro[1]].reshape(L, 10, Wd)
    feat = feat.cpu().numpy()[:fo[1]].reshape(L, 10)
    r0, f0 = golden["pca_reduced"], golden["pca_features"]
    assert_pca_parity(red, r0, feat, f0, "golden", clips=c)
