"""GPU parity of the spectral features (hb_welch_psd_batch and
hb_spectral_features_batch, mirroring welch_psd of catch22/fft.rs:120-170 and
spectral_features / extract_spectral_features of spectral.rs:20-151) against
the oracle's numpy restatement.

The GPU forms each segment's DFT as a product with an FP64 twiddle table on
the matrix cores; the oracle uses numpy's FFT. PSD bins agree to rtol 1e-9
plus atol 1e-12 x the row maximum (bins near zero carry the transform's
absolute rounding); the f32 features to rtol 1e-5, atol 1e-5.
"""
import numpy as np
import pytest

from hyburn_amd.features import (extract_spectral_features, spectral_batch, spectral_features, welch_batch,
                                 welch_psd)

pytestmark = pytest.mark.gpu


def _sine(n, f, fs):
    return np.sin(np.arange(n, dtype=np.float64) * 2.0 * np.pi * f / fs).astype(np.float32)


def _check_psd(P, R):
    atol = 1e-12 * max(np.abs(R).max(), 1e-300)
    np.testing.assert_allclose(P, R, rtol=1e-9, atol=atol)


@pytest.mark.parametrize("T,nfft,w", [(23, 20, 6), (64, 64, 64), (100, 0, 16), (1200, 2048, 256), (10, 16, 12),
                                      (257, 300, 33)])
def test_welch_matches_oracle(gpu, orc, T, nfft, w):
    rng = np.random.default_rng(T + w)
    X = np.cumsum(rng.normal(size=(3, T, 5)), axis=1).astype(np.float32)
    P, freqs = welch_batch(X, nfft, 250.0, w)
    for b in range(3):
        for n in range(5):
            R, rf = orc.welch_psd(X[b, :, n].astype(np.float64), nfft, 250.0, w)
            _check_psd(P[b, n], R)
    np.testing.assert_allclose(freqs, rf, rtol=1e-15)


def test_welch_closed_forms(gpu):
    c = np.cos(2 * np.pi * 5 * np.arange(64) / 64).astype(np.float32)
    psd, _ = welch_psd(c, 64, 1.0, 64)
    # f32 input: the cosine's rounding leaves ~1e-7 relative leakage
    assert abs(psd[5] * 64 - 32.0 ** 2) < 1e-3 and np.all(np.abs(np.delete(psd, 5)) < 1e-9)
    y = np.random.default_rng(5).normal(size=64).astype(np.float32)
    psd, _ = welch_psd(y, 64, 1.0, 64)
    d = y.astype(np.float64) - y.astype(np.float64).mean()
    np.testing.assert_allclose(psd[0] + 2 * psd[1:-1].sum() + psd[-1], (d * d).sum(), rtol=1e-12)


@pytest.mark.parametrize("T", [16, 32, 100, 1024, 1200, 4096])
def test_spectral_features_match_oracle(gpu, orc, T):
    rng = np.random.default_rng(T)
    X = np.cumsum(rng.normal(size=(2, T, 7)), axis=1).astype(np.float32)
    X[0, :, 3] = 2.5                                   # constant: zero power -> NaN
    F = spectral_batch(X, 1000.0)
    assert F.shape == (2, 7, 9)
    for b in range(2):
        for n in range(7):
            np.testing.assert_allclose(F[b, n], orc.spectral_features(X[b, :, n], 1000.0), rtol=1e-5, atol=1e-5)
    assert np.all(np.isnan(F[0, 3]))


def test_spectral_reference_kats(gpu):
    # spectral.rs:157-264
    F = spectral_features(_sine(1024, 10.0, 100.0), 100.0)
    assert F[2] > 0.5 and abs(F[5] - 10.0) < 3.0
    F = spectral_features(_sine(1024, 6.0, 100.0), 100.0)
    assert F[1] > 0.5 and F[0] < 0.3
    assert spectral_features(_sine(2048, 40.0, 200.0), 200.0)[4] > 0.5
    assert np.all(np.isnan(spectral_features(np.ones(4, np.float32), 100.0)))
    assert not np.any(np.isinf(spectral_features(_sine(32, 10.0, 100.0), 100.0)))
    traj = np.sin(np.arange(64 * 2 * 3, dtype=np.float32) * 0.1)
    assert extract_spectral_features(traj, (64, 2, 3, 1), 100.0).shape == (18,)


def test_extract_spectral_features_matches_oracle(gpu, orc):
    rng = np.random.default_rng(9)
    T, nvar, nn, nm = 500, 2, 6, 2
    traj = np.cumsum(rng.normal(size=(T, nvar, nn, nm)), axis=0).astype(np.float32)
    got = extract_spectral_features(traj.ravel(), (T, nvar, nn, nm), 1000.0)
    ref = []
    for v in range(nvar):
        acc = np.zeros(9)
        for n in range(nn):
            for m in range(nm):
                acc += orc.spectral_features(traj[:, v, n, m], 1000.0).astype(np.float64)
        ref.append((acc / (nn * nm)).astype(np.float32))
    np.testing.assert_allclose(got, np.concatenate(ref), rtol=1e-5, atol=1e-5)


def test_spectral_full_batch(gpu, orc):
    # 64 series x 1200 samples x 84 nodes in one call; spot checks plus the
    # invariants: band fractions in [0, 1] summing to <= 1, spread >= 0
    rng = np.random.default_rng(2)
    X = (np.cumsum(rng.normal(size=(64, 1200, 84)), axis=1) + rng.normal(size=(64, 1200, 84))).astype(np.float32)
    F = spectral_batch(X, 1000.0)
    for b, n in [(0, 0), (31, 40), (63, 83)]:
        np.testing.assert_allclose(F[b, n], orc.spectral_features(X[b, :, n], 1000.0), rtol=1e-5, atol=1e-5)
    assert np.all((F[..., :5] >= 0) & (F[..., :5] <= 1)) and np.all(F[..., :5].sum(-1) <= 1 + 1e-5)
    assert np.all(F[..., 6] >= 0)
