"""Functional-connectivity features on the GPU, mirroring src/sbi/features/fc.rs.

  fc_matrix(trajectory)                  fc.rs:11-57    one [T][N] series -> [N][N]
  fcd_matrix(trajectory, window_len, stride)
                                         fc.rs:94-179   one series -> [W][W] (empty when W <= 1)
  fc_batch(series) / fcd_batch(series, ...)
                                         the same over [B][T][N] in one call

The reference takes a flat slice plus n_steps / nnodes; here the shape comes
from the array. Invalid windows raise HyburnError (InvalidConfig) where
fcd_matrix asserts. The matrices are computed by the HIP library; there is no
CPU path.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L


def _series(x) -> np.ndarray:
    a = np.ascontiguousarray(x, np.float32)
    if a.ndim == 2:
        a = a[None]
    if a.ndim != 3:
        raise ValueError("series must be [n_steps, n_nodes] or [n_series, n_steps, n_nodes]")
    return a


def fcd_windows(n_steps: int, window_len: int, stride: int) -> int:
    """(n_steps - window_len) // stride + 1 (fc.rs:104); 0 for an invalid window."""
    return int(L.lib().hb_fcd_windows(n_steps, window_len, stride))


def fc_batch(series, device: int = 0) -> np.ndarray:
    """Pearson FC of every series: [B][T][N] -> [B][N][N] float32."""
    a = _series(series)
    B, T, N = a.shape
    out = np.empty((B, N, N), np.float32)
    L.check(L.lib().hb_fc_batch(a.ctypes.data_as(L.fp), B, T, N, out.ctypes.data_as(L.fp), device))
    return out


def fcd_batch(series, window_len: int, stride: int, device: int = 0) -> np.ndarray:
    """FCD of every series: [B][T][N] -> [B][W][W] float32 (W = 0 when W <= 1)."""
    a = _series(series)
    B, T, N = a.shape
    W = fcd_windows(T, window_len, stride)
    W = W if W > 1 else 0
    out = np.empty((B, W, W), np.float32)
    L.check(L.lib().hb_fcd_batch(a.ctypes.data_as(L.fp), B, T, N, window_len, stride,
                                 out.ctypes.data_as(L.fp), device))
    return out


def fc_matrix(trajectory, device: int = 0) -> np.ndarray:
    """FC of one [T][N] series (fc.rs:11-57)."""
    return fc_batch(np.asarray(trajectory)[None], device)[0]


def fcd_matrix(trajectory, window_len: int, stride: int, device: int = 0) -> np.ndarray:
    """FCD of one [T][N] series (fc.rs:94-179)."""
    return fcd_batch(np.asarray(trajectory)[None], window_len, stride, device)[0]
