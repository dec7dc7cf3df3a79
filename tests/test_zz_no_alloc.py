"""hb_fc_batch_noalloc: the FC entry point that allocates nothing.

It takes device pointers and a stream, and keeps its jobs, column sums and
split partials in a fixed 64 MiB device array per GPU. Its kernels and
launch shapes are those of hb_fc_batch, so its result is compared with
hb_fc_batch bit for bit, on two calls in a row (the second smaller, so the
scratch holds the first call's contents when the second starts), and sizes
beyond the scratch are refused before any GPU work."""
import ctypes as C

import numpy as np
import pytest
import torch

from hyburn_amd import _lib as L

pytestmark = pytest.mark.gpu


def _lib():
    return L.lib()


def _series(B, T, N, seed):
    rng = np.random.default_rng(seed)
    X = np.cumsum(rng.normal(size=(B, T, N)), axis=1) + np.cumsum(rng.normal(size=(B, T, 1)), axis=1)
    return np.ascontiguousarray(X, np.float32)


def _noalloc(lib, X):
    B, T, N = X.shape
    dx = torch.from_numpy(X).to("cuda:0")
    dfc = torch.full((B, N, N), float("nan"), dtype=torch.float32, device="cuda:0")
    stream = torch.cuda.current_stream(0)
    rc = lib.hb_fc_batch_noalloc(C.c_void_p(dx.data_ptr()), B, T, N, C.c_void_p(dfc.data_ptr()), 0,
                                 C.c_void_p(stream.cuda_stream))
    assert rc == L.HB_OK, lib.hb_last_error().decode()
    stream.synchronize()
    return dfc.cpu().numpy()


def _alloc(lib, X):
    from hyburn_amd.features import fc_batch
    return fc_batch(X)


@pytest.mark.parametrize("shape", [((48, 1200, 84), (5, 300, 20)), ((3, 700, 150), (2, 40, 7))])
def test_noalloc_fc_matches_allocating_entry_point(gpu, shape):
    lib = _lib()
    for k, (B, T, N) in enumerate(shape):
        print(f"[noalloc] call {k}: {B} x {T} x {N}", flush=True)
        X = _series(B, T, N, 7 + k)
        got = _noalloc(lib, X)
        want = _alloc(lib, X)
        np.testing.assert_array_equal(got, want)


def test_noalloc_fc_refuses_sizes_beyond_its_scratch(gpu):
    lib = _lib()
    # 500 series of 1200 samples x 84 nodes: one 96-column block pair per series,
    # 500 blocks split in 2 sample ranges -> 1000 FP64 partials of 75,264 B
    # (75 MB) beyond the 64 MiB scratch; refused before any GPU work
    dx = torch.zeros(1, dtype=torch.float32, device="cuda:0")
    rc = lib.hb_fc_batch_noalloc(C.c_void_p(dx.data_ptr()), 500, 1200, 84, C.c_void_p(dx.data_ptr()), 0, None)
    assert rc == L.HB_INVALID_CONFIG and b"scratch" in lib.hb_last_error()


def test_noalloc_fc_refuses_host_pointers(gpu):
    lib = _lib()
    X = _series(2, 64, 5, 1)
    out = np.zeros((2, 5, 5), np.float32)
    rc = lib.hb_fc_batch_noalloc(X.ctypes.data_as(C.c_void_p), 2, 64, 5, out.ctypes.data_as(C.c_void_p), 0, None)
    assert rc == L.HB_INVALID_CONFIG and b"device memory" in lib.hb_last_error()
