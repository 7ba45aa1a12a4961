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


# Bounds on the PCA features, as max |x - ref| relative to max |ref| over
# all components of a neighborhood.  The reference is the FP64 Gram form of
# the reference's two-stage PCA (oracle.core.*_fp64).  Against the oracle's
# float form (the reference's own arithmetic) the bound adds the oracle's
# own distance from FP64: the later power-method components of some
# neighborhoods are sensitive enough that FP32-level arithmetic moves them
# by ~1e-3, in the reference as in any implementation.
PCA_STAGE1_BOUND = 1e-3
PCA_STAGE2_BOUND = 5e-4


def _rel(a, b):
    import numpy as np
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / max(float(np.max(np.abs(b))), 1e-30))


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


def assert_pca_parity(r_g, r_o, f_g, f_o, label="", clips=None, fp64=None):
    """Assert every stage-1 component and stage-2 feature of the GPU (r_g,
    f_g) against the FP64 Gram form within the bounds above, and against the
    oracle's float form (r_o, f_o) within the bounds plus the oracle's own
    distance from FP64.  fp64: (r_64, f_64), or computed from clips
    ((L, T, W)).  Returns (e1, e2) against FP64."""
    from oracle import core as oc
    if fp64 is None:
        r64 = oc.dimension_reduce_clips_fp64(clips, r_o.shape[1], 10000)
        f64 = oc.stage2_features_fp64(r64, f_o.shape[1], 10000)
    else:
        r64, f64 = fp64
    e1, per1, e2, per2 = pca_errors(r_g, r64, f_g, f64)
    o1, o2 = _rel(r_o, r64), _rel(f_o, f64)
    g1, g2 = _rel(r_g, r_o), _rel(f_g, f_o)
    say(f"  {label} vs FP64: stage1 {e1:.2e} stage2 {e2:.2e} (per feature "
        f"{' '.join(f'{v:.0e}' for v in per2)}); vs oracle: {g1:.2e} {g2:.2e}; "
        f"oracle vs FP64: {o1:.2e} {o2:.2e}")
    assert e1 < PCA_STAGE1_BOUND, f"{label}: stage-1 error {e1:.3e} against FP64"
    assert e2 < PCA_STAGE2_BOUND, f"{label}: stage-2 error {e2:.3e} against FP64"
    assert g1 < PCA_STAGE1_BOUND + o1, f"{label}: stage-1 distance {g1:.3e} from the oracle"
    assert g2 < PCA_STAGE2_BOUND + o2, f"{label}: stage-2 distance {g2:.3e} from the oracle"
    return e1, e2
