"""CPU checks of the boundary: the C-ABI library loads and exports every entry
point include/hyburn_hip.h declares; host-side validation returns the
reference's error classes without touching a GPU; the Philox noise stream
matches the published Random123 known-answer vectors."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from hyburn_amd import _lib as L
from hyburn_amd.instances import Instance, _Marshal, _opts
from hyburn_amd.synth import G2DO_DEFAULT, _philox4x32_10, philox_normal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "hyburn_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hb_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = _declared_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), f"{n} declared but not exported"
        assert n in L.PROTOTYPES, f"{n} has no ctypes prototype"


def test_struct_sizes_match_header():
    # natural-alignment layout of the C structs, computed independently
    assert C.sizeof(L.HbRunOpts) == 16
    assert C.sizeof(L.HbRunStats) == 40
    assert C.sizeof(L.HbSweepPoint) == 32
    assert C.sizeof(L.HbAdjointSpec) == 64
    assert C.sizeof(L.HbAdjointOpts) == 12
    assert C.sizeof(L.HbAdjointStats) == 32


def test_version_and_backend_selection():
    lib = L.lib()
    assert b"gfx950" in lib.hb_version()
    assert lib.hb_select_backend(b"hip") == b"hip"
    assert lib.hb_select_backend(b"rocm") == b"hip"
    assert lib.hb_select_backend(b"ndarray") is None
    assert b"HIP only" in lib.hb_last_error()


def _plan_error(inst):
    m = _Marshal([inst])
    h = C.c_void_p()
    o = _opts(0, False, True, 0, 0)
    rc = L.lib().hb_plan_create(m.arr, 1, C.byref(o), C.byref(h))
    return rc, L.lib().hb_last_error().decode()


def _base(**kw):
    d = dict(model=L.G2DO, n_nodes=4, n_steps=10, dt=0.1, params=G2DO_DEFAULT, x0=np.zeros((4, 2), np.float32))
    d.update(kw)
    return Instance(**d)


@pytest.mark.parametrize("kw,needle", [
    (dict(dt=0.0), "dt must be positive"),
    (dict(n_steps=-1), "n_steps"),
    (dict(wkind=L.W_CSR, indptr=np.array([0, 1, 1, 1, 1], np.int32), indices=np.array([9], np.int32),
          weights=np.array([1.0], np.float32)), "CSR index out of range"),
    (dict(wkind=L.W_CSR, indptr=np.array([0, 1, 1, 1, 2], np.int32), indices=np.array([1], np.int32),
          weights=np.array([1.0], np.float32)), "indptr inconsistent"),
    (dict(integrator=L.HEUN_STOCH), "nsig"),
    (dict(cfn=L.CF_KURAMOTO, wkind=L.W_DENSE, dense=np.zeros((4, 4), np.float32)), "Kuramoto coupling"),
    (dict(model=L.MPR, params=np.ones(7, np.float32), cfn=L.CF_SJR), "combination"),
    # PreSigmoidal's dynamic threshold (coupling_params[5]) has no valid
    # configuration in the reference (construction.rs:731-740, dfun.rs:451-452)
    (dict(model=L.MPR, params=np.ones(7, np.float32), cfn=L.CF_PRESIGMOIDAL,
          cparams=(1.0, 1.0, 1.0, 1.0, 0.5, 1.0), wkind=L.W_DENSE, dense=np.zeros((4, 4), np.float32)),
     "dynamic threshold"),
    (dict(cfn=L.CF_PRESIGMOIDAL, cparams=(1.0, 1.0, 1.0, 1.0, 0.5, 1.0, 1.0), wkind=L.W_DENSE,
          dense=np.zeros((4, 4), np.float32)), "dynamic threshold"),
])
def test_plan_validation_errors(kw, needle):
    rc, msg = _plan_error(_base(**kw))
    assert rc == L.HB_INVALID_CONFIG and needle in msg


def test_engine_create_validation():
    lib = L.lib()
    cfg = L.HbSimConfig()
    cfg.dt, cfg.sim_length = 0.0, 10.0
    h = C.c_void_p()
    assert lib.hb_engine_create(C.byref(cfg), 1, 0, C.byref(h)) == L.HB_INVALID_CONFIG
    assert b"dt must be positive" in lib.hb_last_error()
    cfg.dt = 0.1
    assert lib.hb_engine_create(C.byref(cfg), 1, 0, C.byref(h)) == L.HB_INVALID_CONFIG
    assert b"subnetwork" in lib.hb_last_error()
    p = np.asarray(G2DO_DEFAULT[:5], np.float32)
    x0 = np.zeros(8, np.float32)
    sub = L.HbSubnetwork(L.G2DO, 4, 5, p.ctypes.data_as(L.fp), x0.ctypes.data_as(L.fp))
    cfg.subnetworks, cfg.n_subnetworks = C.pointer(sub), 1
    assert lib.hb_engine_create(C.byref(cfg), 1, 0, C.byref(h)) == L.HB_INVALID_CONFIG
    assert b"parameter count" in lib.hb_last_error()


# Random123 known-answer vectors for philox4x32-10 (kat_vectors).
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    c = [np.array([v], np.uint32) for v in ctr]
    out = _philox4x32_10(*c, key[0], key[1])
    assert tuple(int(o[0]) for o in out) == want


def test_philox_normal_moments_and_determinism():
    z = philox_normal(123, 5, 400_001)
    assert z.dtype == np.float32 and z.shape == (400_001,)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    np.testing.assert_array_equal(z[:1000], philox_normal(123, 5, 1000))
    assert not np.array_equal(z[:1000], philox_normal(123, 6, 1000))


def _toml_g2do(n=4):
    x0 = ", ".join(["0.0"] * (2 * n))
    return f"""
sim_length = 10.0
dt = 0.1
integrator = "heun"
[network]
[[network.subnetworks]]
model = "Generic2dOscillator"
nnodes = {n}
params = [1, 0, -2, -10, 0, 0.02, 3, 1, 0, 1, 1, 1]
initial_state = [{x0}]
[[network.projections]]
src = 0
tgt = 0
weights = 0.01
coupling_fn = "Linear"
coupling_params = [0.01]
"""


def test_engine_rejects_value_count_mismatch():
    """n_values != n_sweep is an InvalidConfig error (checked before any device
    work); the reference broadcasts it unchecked."""
    from hyburn_amd.engine import BatchHybridEngine, SimConfig, SweepParam
    cfg = SimConfig.from_toml_str(_toml_g2do())
    assert cfg.n_steps() == 100
    eng = BatchHybridEngine.from_config(cfg, n_sweep=3)
    with pytest.raises(L.HyburnError) as e:
        eng.run_sweep(SweepParam(0, 1), [0.0, 1.0], 10)
    assert e.value.code == L.HB_INVALID_CONFIG
    eng.close()


def test_config_parse_errors():
    from hyburn_amd.engine import SimConfig, parse_sweep_param
    with pytest.raises(L.HyburnError) as e:
        SimConfig.from_toml_str("sim_length = ")
    assert e.value.code == L.HB_INVALID_CONFIG
    with pytest.raises(L.HyburnError) as e:
        SimConfig.from_file("/nonexistent/cfg.toml")
    assert e.value.code == L.HB_IO_ERROR
    cfg = SimConfig.from_toml_str(_toml_g2do())
    assert parse_sweep_param("subnetworks[0].params[2]", cfg) == (0, 2)
    with pytest.raises(L.HyburnError):
        parse_sweep_param("dt", cfg)


def test_n_steps_f64_truncation():
    """(sim_length / dt) as usize truncates in f64 (cli.rs:901): 4999.95 ms at
    dt = 0.05 gives 99998 steps, not 99999."""
    from hyburn_amd.engine import SimConfig
    cfg = SimConfig.from_toml_str(_toml_g2do())
    cfg.sim_length, cfg.dt = 4999.95, 0.05
    assert cfg.n_steps() == 99998


def test_write_npy_f32_format(tmp_path):
    """v1.0 header, '<f4', total header length a multiple of 64 (io.rs:88-125);
    numpy reads the file back with the written shape."""
    from hyburn_amd.engine import write_npy_f32
    data = np.arange(24, dtype=np.float32)
    p = tmp_path / "x.npy"
    write_npy_f32(str(p), data, [2, 3, 4])
    raw = p.read_bytes()
    assert raw[:8] == b"\x93NUMPY\x01\x00"
    hlen = int.from_bytes(raw[8:10], "little")
    assert (10 + hlen) % 64 == 0 and raw[10 + hlen - 1:10 + hlen] == b"\n"
    back = np.load(str(p), allow_pickle=False)
    assert back.shape == (2, 3, 4) and back.dtype == np.dtype("<f4")
    np.testing.assert_array_equal(back.ravel(), data)


def test_fc_fcd_host_validation():
    # fcd_matrix asserts (fc.rs:101-102) -> InvalidConfig before any device call
    lib = L.lib()
    x = np.zeros((2, 10, 3), np.float32)
    out = np.zeros(1000, np.float32)
    xp, op = x.ctypes.data_as(L.fp), out.ctypes.data_as(L.fp)
    assert lib.hb_fcd_windows(10, 4, 2) == 4
    assert lib.hb_fcd_windows(10, 10, 3) == 1
    assert lib.hb_fcd_windows(10, 1, 1) == 0 and lib.hb_fcd_windows(10, 11, 1) == 0
    assert lib.hb_fcd_windows(10, 4, 0) == 0
    for L_, S_ in [(1, 1), (11, 1), (4, 0)]:
        assert lib.hb_fcd_batch(xp, 2, 10, 3, L_, S_, op, 0) == L.HB_INVALID_CONFIG
    assert lib.hb_fcd_batch(xp, 0, 10, 3, 4, 2, op, 0) == L.HB_INVALID_CONFIG
    assert lib.hb_fc_batch(xp, 2, 0, 3, op, 0) == L.HB_INVALID_CONFIG
    assert lib.hb_fc_batch(None, 2, 10, 3, op, 0) == L.HB_INVALID_CONFIG
    # a single window: the reference returns an empty matrix, nothing is written
    out[:] = 7.0
    assert lib.hb_fcd_batch(xp, 2, 10, 3, 10, 1, op, 0) == L.HB_OK
    assert np.all(out == 7.0)


def test_welch_host_validation():
    # welch_psd divides by window_width / 2 (catch22/fft.rs:126-131): width < 2 is rejected
    lib = L.lib()
    x = np.zeros((1, 40, 2), np.float32)
    xp = x.ctypes.data_as(L.fp)
    psd = np.zeros(4096)
    pp = psd.ctypes.data_as(C.POINTER(C.c_double))
    assert lib.hb_welch_bins(0, 16) == 9 and lib.hb_welch_bins(100, 16) == 65
    assert lib.hb_welch_bins(64, 1) == 0 and lib.hb_welch_bins(70000, 16) == 0
    assert lib.hb_welch_psd_batch(xp, 1, 40, 2, 64, 100.0, 1, pp, 0) == L.HB_INVALID_CONFIG
    assert lib.hb_welch_psd_batch(xp, 1, 40, 2, 1 << 20, 100.0, 16, pp, 0) == L.HB_INVALID_CONFIG
    # fewer than 16 samples: NaN features without any device work (spectral.rs:21-24)
    f = np.zeros((1, 2, 9), np.float32)
    y = np.ones((1, 10, 2), np.float32)
    assert lib.hb_spectral_features_batch(y.ctypes.data_as(L.fp), 1, 10, 2, 100.0, f.ctypes.data_as(L.fp), 0) == 0
    assert np.all(np.isnan(f))


def test_history_floats():
    """hb_history_floats: max_delay steps x nodes x ring channels (Kuramoto 2
    per cvar; MPR 1 while its cv coefficient is 0, else 2); 0 without delays."""
    from hyburn_amd.instances import _Marshal
    from hyburn_amd.synth import MPR_DEFAULT, KUR_DEFAULT
    ip = np.array([0, 1, 2, 3, 4, 5], np.int32)
    ix = np.array([1, 2, 3, 4, 0], np.int32)
    w = np.ones(5, np.float32)
    d = np.array([3, 12, 1, 7, 9], np.int32)
    csr = dict(wkind=L.W_CSR, indptr=ip, indices=ix, weights=w, edge_delays=d, n_nodes=5,
               x0=np.zeros((5, 2), np.float32))
    cases = [(_base(**csr), 12 * 5),
             (_base(**{**csr, "model": L.MPR, "params": MPR_DEFAULT}), 12 * 5),
             (_base(**{**csr, "model": L.MPR, "params": np.array([1, 1, -5, 15, 0, 1, 0.5], np.float32)}), 12 * 5 * 2),
             (_base(**{**csr, "model": L.KURAMOTO, "params": KUR_DEFAULT, "x0": np.zeros((5, 1), np.float32),
                       "cfn": L.CF_KURAMOTO}), 12 * 5 * 2),
             (_base(), 0)]
    for I, want in cases:
        m = _Marshal([I])
        assert L.lib().hb_history_floats(m.arr) == want


@pytest.mark.parametrize("tr_ms", [0.0, 1.0, 50.0, 0.7, 333.3])
def test_bold_downsample_matches_the_monitor(orc, tr_ms):
    """hb_bold_downsample (host code, no device) resamples a monitor's native
    series as BoldMonitor::downsample_to_tr does (bold_monitor.rs:123-203):
    against the oracle's restatement, bit for bit, over the times of two runs
    with partial windows (window 1 ms); tr 0 or equal to the window returns the
    native samples; a short output buffer is refused."""
    rng = np.random.default_rng(5)
    N, dtb = 7, 1.0
    times = [k * dtb + dtb for k in range(200)] + [200 * dtb + 0.5]
    times += [len(times) + k + 1.0 * dtb for k in range(199)] + [400 * dtb + 0.5]
    tm = np.ascontiguousarray(times, np.float64)
    nat = rng.standard_normal((len(tm), N)).astype(np.float32)
    dp = C.POINTER(C.c_double)
    lib = L.lib()
    rows = lib.hb_bold_downsample(tm.ctypes.data_as(dp), len(tm), nat.ctypes.data_as(L.fp), N, dtb, tr_ms, None, 0)
    got = np.zeros((rows, N), np.float32)
    assert lib.hb_bold_downsample(tm.ctypes.data_as(dp), len(tm), nat.ctypes.data_as(L.fp), N, dtb, tr_ms,
                                  got.ctypes.data_as(L.fp), rows) == rows
    cap = len(tm) * 2
    want = np.zeros((cap, N), np.float32)
    m = orc.lib().hb_orc_bold_downsample(tm.ctypes.data_as(dp), orc._p(nat), len(tm), N, dtb, tr_ms,
                                         orc._p(want), cap)
    assert m == rows
    np.testing.assert_array_equal(got, want[:m])
    if tr_ms in (0.0, dtb):
        np.testing.assert_array_equal(got, nat)
    else:
        assert rows == int(np.floor(tm[-1] / tr_ms))
    if rows > 1:
        assert lib.hb_bold_downsample(tm.ctypes.data_as(dp), len(tm), nat.ctypes.data_as(L.fp), N, dtb, tr_ms,
                                      got.ctypes.data_as(L.fp), rows - 1) == -1
