"""Comparison helpers shared by the tests."""
import numpy as np


def bits(a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.float64:
        return a.view(np.uint64)
    return a


def bit_equal(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(bits(a), bits(b))


def spmv_scaled_error(y, y_ref, rows, cols, vals, x, nrows):
    """max_i |y_i - y_ref_i| / (|A| |x|)_i  (the norm-scaled SpMV error)."""
    scale = np.bincount(rows, weights=np.abs(vals) * np.abs(x[cols]),
                        minlength=nrows)
    scale = np.where(scale == 0, 1.0, scale)
    return float(np.max(np.abs(y - y_ref) / scale)) if nrows else 0.0
