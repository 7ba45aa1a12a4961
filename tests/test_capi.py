"""The C ABI library loads and exports every symbol declared in include/msamd.h;
host-only entry points agree with the oracle.  CPU only (no kernel launches)."""
import ctypes
import os
import re

import pytest

from mountainsort_amd import _capi
from oracle import core as oc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "msamd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(msamd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("libmsamd.so not built")
    L = ctypes.CDLL(_capi.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # every declared symbol has a ctypes signature in the binding
    assert set(names) <= set(_capi.SIGNATURES), set(names) - set(_capi.SIGNATURES)


def test_host_entry_points():
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("libmsamd.so not built")
    L = _capi.lib()
    assert L.msamd_version() >= 10000
    for M, N, T in [(384, 9_000_000, 8), (64, 18_000_000, 8), (64, 18_000_000, 1), (4, 1_800_000, 8),
                    (4, 5000, 8), (1, 100, 3)]:
        assert L.msamd_detect_chunk_size(M, N, T) == oc.detect_chunk_size(M, N, T)
    assert L.msamd_detect_workspace_size(9_000_000, 20000, 384, 10.0) > 0
    assert L.msamd_whiten_covariance_workspace_size(384, 9_000_000) > 0
    assert L.msamd_whitening_matrix_workspace_size(384) >= 384 * 384 * 8
    assert L.msamd_whiten_apply_workspace_size(384) > 0
    assert L.msamd_mask_workspace_size(384, 9_000_000, 2000) > 0
    assert L.msamd_mask_workspace_size(0, 10, 10) < 0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_capi, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_capi, "_lib", None)
    with pytest.raises(_capi.MsamdLibraryMissing):
        _capi.lib()


def test_bandpass_host_entry_points():
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("libmsamd.so not built")
    import numpy as np
    L = _capi.lib()
    for M, T in [(64, 8), (64, 2), (4, 8), (384, 1)]:
        assert L.msamd_bandpass_chunk_size(M, T) == oc.bandpass_chunk_size(M, T)
    H = np.empty(240000)
    assert L.msamd_bandpass_kernel(240000, 30000.0, 300.0, 6000.0, 1000.0,
                                   H.ctypes.data_as(ctypes.c_void_p)) == 0
    np.testing.assert_array_equal(H, oc.define_kernel(240000, 30000, 300, 6000, 1000))
    assert L.msamd_bandpass_supported(120000, 60000) == 1
    # any other window length goes through the chirp-z form, up to 2^25 points
    assert L.msamd_bandpass_supported(1278101, 60000) == 1
    assert L.msamd_bandpass_workspace_size(4, 1_800_000, 1278101, 60000) > 0
    assert L.msamd_bandpass_supported((1 << 25) - 120000, 60000) == 1
    assert L.msamd_bandpass_supported((1 << 25) - 120000 + 2, 60000) == 0
    assert L.msamd_bandpass_workspace_size(384, 9_000_000, 120000, 60000) > 0


def test_pca_workspace_query():
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("libmsamd.so not built")
    import numpy as np
    L = _capi.lib()
    ev = np.array([300, 0, 12345, 50, 1500], np.int64)
    w = np.array([5, 3, 12, 1, 16], np.int32)
    args = (ev.ctypes.data_as(ctypes.c_void_p), w.ctypes.data_as(ctypes.c_void_p))
    assert L.msamd_pca_workspace_size(5, *args, 50, 10, 10, 10000) > 0
    # neighborhoods up to 128 channels (the ms3alg default adjacency_radius = -1
    # puts every channel of a 64-channel probe in each), not beyond
    w_wide = np.array([5, 64, 128, 1, 17], np.int32)
    assert L.msamd_pca_workspace_size(5, args[0], w_wide.ctypes.data_as(ctypes.c_void_p),
                                      50, 10, 10, 10000) > 0
    w_bad = np.array([5, 3, 129, 1, 16], np.int32)
    assert L.msamd_pca_workspace_size(5, args[0], w_bad.ctypes.data_as(ctypes.c_void_p),
                                      50, 10, 10, 10000) < 0
    assert L.msamd_pca_workspace_size(5, *args, 65, 10, 10, 10000) < 0


def test_pca_plan_deferred_create_needs_attach():
    """A plan created without a workspace holds its host tables only: it
    reports the workspace it needs (the same as the size query), and runs
    are refused until a workspace is attached.  Host-only calls."""
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("libmsamd.so not built")
    import numpy as np
    L = _capi.lib()
    ne = np.array([12000, 300, 0, 45000], dtype=np.int64)
    w = np.array([5, 17, 3, 12], dtype=np.int32)
    tiles = np.array([2560, 32, 16], dtype=np.int32)
    args = (4, ne.ctypes.data_as(ctypes.c_void_p), w.ctypes.data_as(ctypes.c_void_p), 50, 10, 10,
            10000, tiles.ctypes.data_as(ctypes.c_void_p))
    nb = L.msamd_pca_plan_workspace_size(*args)
    assert nb > 0
    plan = ctypes.c_void_p()
    assert L.msamd_pca_plan_create(*args, None, 0, None, ctypes.addressof(plan)) == 0
    assert plan.value
    assert L.msamd_pca_plan_workspace_bytes(plan) == nb
    assert L.msamd_pca_plan_run(plan, ctypes.c_void_p(16), ctypes.c_void_p(16),
                                ctypes.c_void_p(16), None) != 0
    assert b"attach" in L.msamd_last_error()
    assert L.msamd_pca_plan_run_timeseries(plan, ctypes.c_void_p(16), 4, 100, ctypes.c_void_p(16),
                                           ctypes.c_void_p(16), ctypes.c_void_p(16),
                                           ctypes.c_void_p(16), ctypes.c_void_p(16), None) != 0
    # attaching a workspace that is too small is refused before any copy
    assert L.msamd_pca_plan_attach(plan, ctypes.c_void_p(256), 1024, None) != 0
    L.msamd_pca_plan_destroy(plan)
