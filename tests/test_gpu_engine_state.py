"""The single-network engine carries its state across runs, as the reference's
BatchHybridEngine does (src/engine/batch_engine/engine.rs:496-540: run_sweep
mutates self.states, self.step, self.histories and self.bold_monitors), and
its batch-averaged BOLD monitor (engine.rs:785-801).

Bars (exact mode, HB_FLAG_EXACT):
* runs of 90 then 70 steps equal one run of 160 steps bit for bit: states,
  trajectories and the second run's tavg (the delay history, step-0 rule and
  batch-window slots cross the boundary), on the batch engine's window
  semantics, on per-edge delays in the one-barrier-per-step kernel and in the
  delay-blocked kernel (HBM and LDS rings);
* the BOLD series a continued engine returns is the whole series of its
  monitors since the start (BoldMonitor::flush at the end of every run,
  bold_monitor.rs:210-227), against the FP32 oracle's Balloon-Windkessel
  step and TR resampling driven by the same windows; the batch-mean monitor
  against the oracle's per-point trajectories averaged over the sweep;
  tolerance: the FP32 resolution of the BOLD signal near rest (one ulp of v
  or q moves it by V0 (k1 + k2) 1.2e-7 = 4.3e-6), 4 such ulps.
"""
import ctypes as C

import numpy as np
import pytest

from hyburn_amd import _lib as L
from hyburn_amd.engine import BatchHybridEngine, MonitorConfig, SweepParam, SweepPoint, SweepSpec

from test_gpu_engine import _csr_net, _engine_cfg, _oracle_inst

pytestmark = pytest.mark.gpu

BOLD_ULP = 4.3e-6


def _split_vs_whole(cfg, n_sweep, vals, param, t1, t2, **kw):
    one = BatchHybridEngine.from_config(cfg, n_sweep, **kw)
    whole = one.run_sweep_with_trajectory(param, vals, t1 + t2, exact=True)
    one.close()
    two = BatchHybridEngine.from_config(cfg, n_sweep, **kw)
    a = two.run_sweep_with_trajectory(param, vals, t1, exact=True)
    b = two.run_sweep_with_trajectory(param, vals, t2, exact=True)
    assert two.step == t1 + t2
    two.close()
    return whole, a, b


@pytest.mark.parametrize("case", ["batch_window", "per_edge_per_step", "per_edge_blocked_lds", "per_edge_blocked_hbm"])
def test_split_run_equals_one_run(gpu, case):
    N, T1, T2 = 16, 90, 70
    kw = {}
    if case == "batch_window":            # the batch engine: delays[0] through the sliding window (Q1)
        ip, ix, w, d = _csr_net(N, 41, max_delay=12)
    elif case == "per_edge_per_step":     # the plain engine's per-edge ring, delays from 1
        ip, ix, w, d = _csr_net(N, 42, max_delay=12)
        kw = dict(engine=L.ENGINE_PLAIN)
    else:                                 # every delay >= 8: the delay-blocked kernel
        rng = np.random.default_rng(43)
        ip, ix, w, _ = _csr_net(N, 43)
        hi = 40 if case.endswith("lds") else 3000   # 3000-step rings exceed LDS
        d = rng.integers(8, hi, len(ix)).astype(np.int32)
        kw = dict(engine=L.ENGINE_PLAIN)
    x0 = np.random.default_rng(44).normal(0, 0.3, (N, 2)).astype(np.float32)
    cfg = _engine_cfg("Generic2dOscillator", N, x0, ip, ix, w, d, T=T1 + T2, cp=(0.2, 0.0))
    vals = np.array([0.0, 0.3, -0.2], np.float32)
    whole, a, b = _split_vs_whole(cfg, 3, vals, SweepParam(0, 1), T1, T2, **kw)
    tw = whole.trajectories[0].reshape(T1 + T2, 3, N, 2)
    np.testing.assert_array_equal(a.trajectories[0].reshape(T1, 3, N, 2), tw[:T1])
    np.testing.assert_array_equal(b.trajectories[0].reshape(T2, 3, N, 2), tw[T1:])
    np.testing.assert_array_equal(b.final_states[0], whole.final_states[0])
    # tavg of the second run: its own steps summed in order, times 1/T2 (engine.rs:813-818)
    acc = np.zeros((3, N, 2), np.float32)
    for t in range(T1, T1 + T2):
        acc = (acc + tw[t]).astype(np.float32)
    np.testing.assert_array_equal(b.tavg[0].reshape(3, N, 2), (acc * np.float32(1.0 / T2)).astype(np.float32))


@pytest.mark.parametrize("case", ["batch_window", "plain_no_delay", "per_edge_blocked_hbm", "stochastic"])
def test_checkpoint_file_continues_bit_for_bit(gpu, case, tmp_path):
    """HybridEngine::checkpoint / resume (runtime.rs:519-742) through a file
    into a freshly built engine: after 90 steps and a checkpoint, 70 steps in
    the new engine equal the 70 the checkpointing engine goes on to run, bit
    for bit, and (without noise, which each run draws afresh from its global
    step) the last 70 of one 160-step run. A single point without delays is
    the reference's own version 2 file; the others carry their delay
    histories (version 4)."""
    from hyburn_amd import checkpoint as ck
    N, T1, T2 = 16, 90, 70
    kw, n_sweep, integ, nsig = {}, 3, "heun", 0.0
    if case == "batch_window":
        ip, ix, w, d = _csr_net(N, 51, max_delay=12)
    elif case == "plain_no_delay":
        ip, ix, w, _ = _csr_net(N, 52)
        d = np.zeros(0, np.int32)
        kw, n_sweep = dict(engine=L.ENGINE_PLAIN), 1
    elif case == "per_edge_blocked_hbm":
        rng = np.random.default_rng(53)
        ip, ix, w, _ = _csr_net(N, 53)
        d = rng.integers(8, 3000, len(ix)).astype(np.int32)
        kw = dict(engine=L.ENGINE_PLAIN)
    else:
        ip, ix, w, d = _csr_net(N, 54, max_delay=6)
        integ, nsig, kw = "heun_stochastic", 0.01, dict(noise_seed=99)
    x0 = np.random.default_rng(55).normal(0, 0.3, (N, 2)).astype(np.float32)
    cfg = _engine_cfg("Generic2dOscillator", N, x0, ip, ix, w, d, T=T1 + T2, cp=(0.2, 0.0), integ=integ, nsig=nsig)
    vals = np.array([0.0, 0.3, -0.2][:n_sweep], np.float32)
    param = SweepParam(0, 1)
    two = BatchHybridEngine.from_config(cfg, n_sweep, **kw)
    two.run_sweep(param, vals, T1, exact=True)
    path = str(tmp_path / "split.ckpt")
    two.checkpoint(path)
    cont = two.run_sweep_with_trajectory(param, vals, T2, exact=True)
    two.close()
    assert ck.read(path)["version"] == (2 if case == "plain_no_delay" else 4)
    three = BatchHybridEngine.from_config(cfg, n_sweep, **kw)
    three.resume(path)
    assert three.step == T1
    b = three.run_sweep_with_trajectory(param, vals, T2, exact=True)
    assert three.step == T1 + T2
    three.close()
    np.testing.assert_array_equal(b.trajectories[0], cont.trajectories[0])
    np.testing.assert_array_equal(b.final_states[0], cont.final_states[0])
    np.testing.assert_array_equal(b.tavg[0], cont.tavg[0])
    if nsig == 0.0:
        one = BatchHybridEngine.from_config(cfg, n_sweep, **kw)
        whole = one.run_sweep_with_trajectory(param, vals, T1 + T2, exact=True)
        one.close()
        tw = whole.trajectories[0].reshape(T1 + T2, n_sweep, N, 2)
        np.testing.assert_array_equal(b.trajectories[0].reshape(T2, n_sweep, N, 2), tw[T1:])


def test_restart_flag_and_step_counter(gpu):
    N = 10
    ip, ix, w, d = _csr_net(N, 45)
    x0 = np.random.default_rng(46).normal(0, 0.3, (N, 2)).astype(np.float32)
    cfg = _engine_cfg("Generic2dOscillator", N, x0, ip, ix, w, d, T=50)
    eng = BatchHybridEngine.from_config(cfg, 2)
    vals = [0.0, 0.1]
    r1 = eng.run_sweep(SweepParam(0, 1), vals, 50, exact=True)
    r2 = eng.run_sweep(SweepParam(0, 1), vals, 50, exact=True)
    assert eng.step == 100
    assert not np.array_equal(r1.final_states[0], r2.final_states[0])
    r3 = eng.run_sweep(SweepParam(0, 1), vals, 50, exact=True, restart=True)
    assert eng.step == 50
    np.testing.assert_array_equal(r3.final_states[0], r1.final_states[0])
    eng.close()


def test_run_sweep_multi_continues_from_point_zero(gpu, orc):
    """run_sweep_multi starts every Cartesian point from sweep point 0 of the
    engine's state at step 0 (engine.rs:886-909) and the engine keeps its
    states afterwards; a run_sweep with another point count is then refused."""
    N = 8
    ip, ix, w, d = _csr_net(N, 47)
    x0 = np.random.default_rng(48).normal(0, 0.3, (N, 2)).astype(np.float32)
    cfg = _engine_cfg("Generic2dOscillator", N, x0, ip, ix, w, d, T=40)
    eng = BatchHybridEngine.from_config(cfg, 2)
    r1 = eng.run_sweep(SweepParam(0, 1), [0.2, 0.4], 40, exact=True)
    start = r1.final_states[0].reshape(2, N, 2)[0]
    vals = [0.1, 0.3, 0.5]
    spec = SweepSpec([SweepPoint(0, 1, vals)])   # G2DO I, the batch-native index
    r2 = eng.run_sweep_multi(spec, 30, exact=True)
    fin = r2.final_states[0].reshape(3, N, 2)
    for k, v in enumerate(vals):
        I = _oracle_inst("Generic2dOscillator", N, start, ip, ix, w, d, engine=L.ENGINE_BATCH,
                         delay_mode=L.DELAY_BATCH_COMPAT, T=30, sweep_val=v, record_traj=False)
        np.testing.assert_array_equal(fin[k], orc.run(I)["final_state"])
    with pytest.raises(L.HyburnError) as e:
        eng.run_sweep(SweepParam(0, 1), [0.2, 0.4], 10)
    assert e.value.code == L.HB_INVALID_STATE
    eng.close()


# ----------------------------------------------------------------- BOLD ----

def _bw_step(orc, z, dt_s, st):
    s, f, v, q = (np.ascontiguousarray(a, np.float32) for a in st)
    zz = np.ascontiguousarray(z, np.float32)
    orc.lib().hb_orc_bold_step_f32(len(zz), orc._p(zz), float(dt_s), orc._p(s), orc._p(f), orc._p(v), orc._p(q))
    sig = np.zeros(len(zz), np.float32)
    orc.lib().hb_orc_bold_signal_f32(len(zz), orc._p(v), orc._p(q), orc._p(sig))
    return sig, (s, f, v, q)


def _monitor(orc, runs, N, P, dt_ms, tr_ms):
    """BoldMonitor over consecutive runs (bold_monitor.rs:72-227): windows of P
    steps of the f32 input, flushed at the end of each run; sample times as
    the monitor forms them; the whole series resampled to TR."""
    st = (np.zeros(N, np.float32), np.ones(N, np.float32), np.ones(N, np.float32), np.ones(N, np.float32))
    times, data = [], []
    dtb = dt_ms * P
    for z in runs:                       # z [T][N] f32 neural input of one run
        acc, cnt = np.zeros(N, np.float32), 0
        for t in range(z.shape[0]):
            acc = (acc + z[t]).astype(np.float32)
            cnt += 1
            if cnt >= P:
                sig, st = _bw_step(orc, (acc / np.float32(cnt)).astype(np.float32), np.float32(dt_ms * cnt / 1000.0), st)
                times.append(len(times) * dtb + dt_ms * cnt)
                data.append(sig)
                acc, cnt = np.zeros(N, np.float32), 0
        if cnt:
            sig, st = _bw_step(orc, (acc / np.float32(cnt)).astype(np.float32), np.float32(dt_ms * cnt / 1000.0), st)
            times.append(len(times) * dtb + dt_ms * cnt)
            data.append(sig)
    tm = np.array(times, np.float64)
    nat = np.ascontiguousarray(np.array(data, np.float32))
    cap = len(times) * 64 + 1
    out = np.zeros((cap, N), np.float32)
    m = orc.lib().hb_orc_bold_downsample(tm.ctypes.data_as(C.POINTER(C.c_double)), orc._p(nat), len(times), N,
                                         dt_ms * P, tr_ms, orc._p(out), cap)
    return out[:m]


@pytest.mark.parametrize("mean", [False, True])
def test_bold_monitors_across_runs(gpu, orc, mean):
    """Two runs of 2005 and 1995 steps (partial windows flushed at each end):
    the second run returns the whole BOLD series of both. Per sweep point, or
    one monitor fed var0 averaged over the points (engine.rs:785-801)."""
    N, P, dt = 10, 10, 0.1
    ip, ix, w, d = _csr_net(N, 49)
    x0 = np.stack([np.full(N, 0.1), np.full(N, -2.0)], 1).astype(np.float32)
    mon = MonitorConfig(monitor_type="bold", tr=0.05, bold_period=P)
    cfg = _engine_cfg("MontbrioPazoRoxin", N, x0, ip, ix, w, d, T=4000, cp=(0.1, 0.0), monitors=[mon])
    vals = np.array([-5.0, -4.6, -4.2], np.float32)
    eng = BatchHybridEngine.from_config(cfg, 3, bold_batch_mean=mean)
    a = eng.run_sweep_with_trajectory(SweepParam(0, 4), vals, 2005, exact=True)
    b = eng.run_sweep_with_trajectory(SweepParam(0, 4), vals, 1995, exact=True)
    eng.close()
    ta = a.trajectories[0].reshape(2005, 3, N, 2)[..., 0]
    tb = b.trajectories[0].reshape(1995, 3, N, 2)[..., 0]
    if mean:   # var0 summed over the points in order, divided by their count (mean_dim)
        def pmean(t):
            s = t[:, 0].copy()
            for k in (1, 2):
                s = (s + t[:, k]).astype(np.float32)
            return (s / np.float32(3)).astype(np.float32)
        runs = [[pmean(ta), pmean(tb)]]
    else:
        runs = [[ta[:, k], tb[:, k]] for k in range(3)]
    want = np.array([_monitor(orc, r, N, P, dt, 50.0) for r in runs])
    got = b.bold[0]
    # 201 + 200 native samples up to 400.5 ms: 8 rows at TR = 50 ms
    assert got.shape == want.shape == (1 if mean else 3, 8, N), (got.shape, want.shape)
    np.testing.assert_allclose(got, want, rtol=0, atol=4 * BOLD_ULP)
    # the first run's series is the start of the second's
    np.testing.assert_array_equal(a.bold[0][:, :a.bold[0].shape[1]], got[:, :a.bold[0].shape[1]])
