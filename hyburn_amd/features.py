"""Functional-connectivity and spectral features on the GPU, mirroring
src/sbi/features/fc.rs, spectral.rs and catch22/fft.rs (welch_psd).

  fc_matrix(trajectory)                  fc.rs:11-57    one [T][N] series -> [N][N]
  fcd_matrix(trajectory, window_len, stride)
                                         fc.rs:94-179   one series -> [W][W] (empty when W <= 1)
  fc_batch(series) / fcd_batch(series, ...)
                                         the same over [B][T][N] in one call
  welch_psd(y, nfft, fs, window_width)   catch22/fft.rs:120-170
  spectral_features(series, fs)          spectral.rs:20-99   one series -> 9 values
  extract_spectral_features(trajectory, shape, fs)
                                         spectral.rs:113-151 [T, nvar, nnodes, nmodes] -> nvar * 9
  welch_batch / spectral_batch           the same over every node series of [B][T][N]

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


SPECTRAL_FEATURE_NAMES = ("delta_power", "theta_power", "alpha_power", "beta_power", "gamma_power",
                          "spectral_centroid", "spectral_spread", "spectral_skewness", "spectral_kurtosis")


def welch_bins(nfft: int, window_width: int) -> int:
    return int(L.lib().hb_welch_bins(nfft, window_width))


def welch_batch(series, nfft: int, fs: float, window_width: int, device: int = 0):
    """Welch PSD of every node series: [B][T][N] -> ([B][N][K] float64, freqs[K])."""
    a = _series(series)
    B, T, N = a.shape
    K = welch_bins(nfft, window_width)
    out = np.empty((B, N, max(K, 0)), np.float64)
    L.check(L.lib().hb_welch_psd_batch(a.ctypes.data_as(L.fp), B, T, N, nfft, fs, window_width,
                                       out.ctypes.data_as(L.C.POINTER(L.C.c_double)), device))
    return out, np.arange(K) * (fs / (2 * (K - 1)))


def welch_psd(y, nfft: int, fs: float, window_width: int, device: int = 0):
    """welch_psd of one series (catch22/fft.rs:120-170) -> (psd, freqs)."""
    psd, freqs = welch_batch(np.asarray(y, np.float32).reshape(1, -1, 1), nfft, fs, window_width, device)
    return psd[0, 0], freqs


def spectral_batch(series, fs: float, device: int = 0) -> np.ndarray:
    """spectral_features of every node series: [B][T][N] -> [B][N][9] float32."""
    a = _series(series)
    B, T, N = a.shape
    out = np.empty((B, N, 9), np.float32)
    L.check(L.lib().hb_spectral_features_batch(a.ctypes.data_as(L.fp), B, T, N, fs, out.ctypes.data_as(L.fp),
                                               device))
    return out


def spectral_features(series, fs: float, device: int = 0) -> np.ndarray:
    """spectral_features of one series (spectral.rs:20-99)."""
    return spectral_batch(np.asarray(series, np.float32).reshape(1, -1, 1), fs, device)[0, 0]


def extract_spectral_features(trajectory, shape, fs: float, device: int = 0) -> np.ndarray:
    """spectral.rs:113-151: trajectory [n_steps, nvar, nnodes, nmodes] -> nvar x 9
    features, each the mean over nodes and modes (f64 sum of the f32 values)."""
    if len(shape) != 4:
        raise ValueError("expected trajectory shape [n_steps, nvar, nnodes, nmodes]")
    T, nvar, nn, nm = shape
    x = np.asarray(trajectory, np.float32).reshape(T, nvar, nn * nm).transpose(1, 0, 2)
    f = spectral_batch(np.ascontiguousarray(x), fs, device)          # [nvar][nn*nm][9]
    return (f.astype(np.float64).sum(axis=1) / (nn * nm)).astype(np.float32).reshape(-1)
