import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libmsamd.so")


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from mountainsort_amd import _capi
    _capi.lib()  # raises MsamdLibraryMissing when the HIP library was not built
    return torch.device("cuda:0")


def say(*args):
    """Print and flush a line naming the case about to run on the GPU."""
    print(*args, flush=True)


# Bounds on the PCA features against the oracle (the reference's float-form
# PCA), as max |gpu - oracle| relative to max |oracle| over all components of
# a neighborhood: split-BF16 Grams and FP32 projections against the
# reference's float data-form power method.
PCA_STAGE1_BOUND = 1e-3
PCA_STAGE2_BOUND = 5e-4


def pca_errors(r_g, r_o, f_g, f_o):
    """(stage-1 error over all components, per-component stage-1 errors,
    stage-2 error over all features, per-feature stage-2 errors).
    r_*: (L, nf1, W) reduced clips; f_*: (L, nf2) features."""
    import numpy as np
    s1 = max(float(np.max(np.abs(r_o))), 1e-30)
    s2 = max(float(np.max(np.abs(f_o))), 1e-30)
    d1 = np.abs(r_g.astype(np.float64) - r_o)
    d2 = np.abs(f_g.astype(np.float64) - f_o)
    per1 = d1.max(axis=(0, 2)) / s1
    per2 = d2.max(axis=0) / s2
    return float(per1.max()), per1, float(per2.max()), per2


def assert_pca_parity(r_g, r_o, f_g, f_o, label=""):
    """Assert every stage-1 component and every stage-2 feature against the
    oracle within the bounds above; returns (e1, e2)."""
    e1, per1, e2, per2 = pca_errors(r_g, r_o, f_g, f_o)
    say(f"  {label} stage1 {e1:.2e} (per component max {per1.max():.1e}, first {per1[0]:.1e}), "
        f"stage2 {e2:.2e} (per feature {' '.join(f'{v:.0e}' for v in per2)})")
    assert e1 < PCA_STAGE1_BOUND, f"{label}: stage-1 error {e1:.3e}"
    assert e2 < PCA_STAGE2_BOUND, f"{label}: stage-2 error {e2:.3e}"
    return e1, e2
