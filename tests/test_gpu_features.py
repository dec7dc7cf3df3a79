"""GPU parity of the functional-connectivity features (hb_fc_batch /
hb_fcd_batch, mirroring fc_matrix and fcd_matrix of src/sbi/features/fc.rs)
against the CPU oracle's restatement.

Both compute in FP64 and round to f32 once; the GPU sums in another order
(MFMA tiles over 32-sample chunks), so FC entries agree to atol 2e-6 and FCD
entries, correlations of those f32 FC values, to atol 2e-5. Sizes cover one
and several 96-column blocks of the correlation matrix (N = 97, 150; FCD with
161 windows).
"""
import numpy as np
import pytest

from hyburn_amd import _lib as L
from hyburn_amd.features import fc_batch, fc_matrix, fcd_batch, fcd_matrix

pytestmark = pytest.mark.gpu

FC_ATOL, FCD_ATOL = 2e-6, 2e-5


def _series(rng, B, T, N):
    # random walks plus a shared component: correlations spread over (-1, 1)
    x = np.cumsum(rng.normal(size=(B, T, N)), axis=1) + np.cumsum(rng.normal(size=(B, T, 1)), axis=1)
    return x.astype(np.float32)


@pytest.mark.parametrize("N", [1, 2, 7, 16, 17, 68, 74, 84, 96, 97, 150])
@pytest.mark.parametrize("T", [1, 3, 33, 400])
def test_fc_batch_matches_oracle(gpu, orc, N, T):
    rng = np.random.default_rng(N * 1000 + T)
    X = _series(rng, 3, T, N)
    F = fc_batch(X)
    assert F.shape == (3, N, N)
    for b in range(3):
        np.testing.assert_allclose(F[b], orc.fc(X[b]), atol=FC_ATOL, rtol=0)
        np.testing.assert_array_equal(F[b], F[b].T)


@pytest.mark.parametrize("N", [5, 120])
def test_fc_special_columns(gpu, orc, N):
    # fc.rs:33-36 constant series -> std 1 -> zero row and column (diagonal too);
    # fc.rs:225-249 perfectly correlated and anti-correlated pairs -> +-1
    rng = np.random.default_rng(7)
    X = rng.normal(size=(300, N)).astype(np.float32)
    X[:, 1] = 2.0 * X[:, 0] + 0.5
    X[:, 2] = -X[:, 0]
    X[:, 3] = 4.0
    F = fc_matrix(X)
    np.testing.assert_allclose(F, orc.fc(X), atol=FC_ATOL, rtol=0)
    assert abs(F[0, 1] - 1.0) < 1e-6 and abs(F[0, 2] + 1.0) < 1e-6
    assert np.all(F[3] == 0.0) and np.all(F[:, 3] == 0.0)
    # the reference's own two-node KATs
    np.testing.assert_allclose(fc_matrix(np.array([[1, 1], [2, 2], [3, 3]], np.float32)), 1.0, atol=1e-5)
    assert abs(fc_matrix(np.array([[1, -1], [2, -2], [3, -3]], np.float32))[0, 1] + 1.0) < 1e-5


@pytest.mark.parametrize("N,T,Lw,S", [(1, 50, 10, 5), (2, 100, 20, 10), (9, 230, 40, 15),
                                      (68, 300, 60, 10), (84, 500, 100, 7), (120, 200, 50, 25),
                                      (10, 500, 20, 3)])
def test_fcd_batch_matches_oracle(gpu, orc, N, T, Lw, S):
    rng = np.random.default_rng(N + T)
    X = _series(rng, 2, T, N)
    D = fcd_batch(X, Lw, S)
    W = (T - Lw) // S + 1
    assert D.shape == (2, W, W)
    for b in range(2):
        np.testing.assert_allclose(D[b], orc.fcd(X[b], Lw, S), atol=FCD_ATOL, rtol=0)
        assert np.all(np.diag(D[b]) == 1.0)
        np.testing.assert_array_equal(D[b], D[b].T)


def test_fcd_edge_cases(gpu):
    X = np.sin(np.arange(30, dtype=np.float32)).reshape(10, 3)
    D = fcd_matrix(X, 4, 2)                       # fc.rs:268-275
    assert D.shape == (4, 4) and abs(D[0, 0] - 1.0) < 1e-5
    assert fcd_matrix(X, 10, 1).shape == (0, 0)    # one window -> empty
    for Lw, S in [(1, 1), (11, 1), (4, 0)]:
        with pytest.raises(L.HyburnError):
            fcd_matrix(X, Lw, S)
    # a series repeating every stride: identical windows -> all ones
    rng = np.random.default_rng(3)
    R = np.tile(rng.normal(size=(20, 6)).astype(np.float32), (8, 1))
    np.testing.assert_allclose(fcd_matrix(R, 40, 20), 1.0, atol=1e-5)


def test_fc_full_sweep_size(gpu, orc):
    # 128 series x 2000 samples x 74 nodes in one call: a spot check against
    # the oracle plus the size-independent properties on every matrix
    rng = np.random.default_rng(5)
    X = _series(rng, 128, 2000, 74)
    F = fc_batch(X)
    for b in (0, 63, 127):
        np.testing.assert_allclose(F[b], orc.fc(X[b]), atol=FC_ATOL, rtol=0)
    assert np.all(np.abs(F) <= 1.0)
    np.testing.assert_array_equal(F, np.transpose(F, (0, 2, 1)))
    np.testing.assert_allclose(np.diagonal(F, axis1=1, axis2=2), 1.0, atol=1e-6)
    # FC is invariant to a per-node affine map with positive scale
    Y = X * rng.uniform(0.5, 2.0, size=(1, 1, 74)).astype(np.float32) + 3.0
    np.testing.assert_allclose(fc_batch(Y[:4]), F[:4], atol=1e-5)
