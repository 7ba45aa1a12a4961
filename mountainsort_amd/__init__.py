"""MI355X-native MountainSort front end.

FFT bandpass filter, artifact masking, whitening, event detection, clip
extraction and two-stage per-neighborhood PCA features, as hand-written gfx950
HIP kernels behind a C ABI (include/msamd.h), with the MountainLab processor
interface of the reference (packages/ms3, packages/mountainsortalg) on top.
"""
from ._capi import LIB_PATH, MsamdError, MsamdLibraryMissing  # noqa: F401

__all__ = ["LIB_PATH", "MsamdError", "MsamdLibraryMissing"]
