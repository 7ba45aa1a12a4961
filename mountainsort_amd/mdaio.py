"""numpy reader and writer of .mda array files.

Format (packages/ms3/mda/mdaio.cpp:50-199): int32 type code, int32 bytes per
entry, int32 number of dimensions (negative: 64-bit dimensions follow), the
dimensions, then the data in column-major order.  A timeseries M x N is
returned as a numpy array of shape (N, M) (row t holds the M channels of
sample t), which is the same memory order.
"""
import numpy as np

_CODES = {-2: np.uint8, -3: np.float32, -4: np.int16, -5: np.int32, -6: np.uint16,
          -7: np.float64, -8: np.uint32}
_TYPES = {np.dtype(v): k for k, v in _CODES.items()}


def readmda(path):
    """Returns the array with its dimensions reversed (numpy C order), so an
    M x N file gives shape (N, M)."""
    with open(path, "rb") as f:
        code, _nbytes, nd = np.frombuffer(f.read(12), dtype="<i4")
        if int(code) not in _CODES:
            raise ValueError(f"{path}: unsupported .mda data type {code}")
        if nd < 0:
            dims = np.frombuffer(f.read(8 * -nd), dtype="<u8").astype(np.int64)
        else:
            dims = np.frombuffer(f.read(4 * nd), dtype="<i4").astype(np.int64)
        dt = np.dtype(_CODES[int(code)])
        n = int(np.prod(dims))
        data = np.fromfile(f, dtype=dt, count=n)
    if data.size != n:
        raise ValueError(f"{path}: truncated data")
    return data.reshape(tuple(int(d) for d in dims[::-1]))


def writemda(path, arr, dims=None):
    """Writes arr; dims (first index fastest) default to arr.shape reversed."""
    arr = np.ascontiguousarray(arr)
    if arr.dtype not in _TYPES:
        raise ValueError(f"unsupported dtype {arr.dtype}")
    if dims is None:
        dims = list(arr.shape[::-1])
    dims = [int(d) for d in dims]
    while len(dims) > 2 and dims[-1] == 1:
        dims.pop()
    while len(dims) < 2:
        dims.append(1)
    big = any(d > 2_000_000_000 for d in dims)
    with open(path, "wb") as f:
        f.write(np.array([_TYPES[arr.dtype], arr.dtype.itemsize,
                          -len(dims) if big else len(dims)], "<i4").tobytes())
        f.write(np.array(dims, "<u8" if big else "<i4").tobytes())
        arr.tofile(f)
