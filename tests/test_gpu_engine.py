"""GPU parity of the config-level API (BatchHybridEngine mirror, hb_engine_* /
hb_run_sweep* in include/hyburn_hip.h) against the CPU oracle.

Each sweep point of the engine is restated as the oracle instance the
reference semantics define (engine.rs:145-471 for the batch engine's weights,
delays and swept dfun form; construction.rs:639-655 for the plain engine), and
compared bit for bit in exact mode (HB_FLAG_EXACT) for G2DO/MPR, whose
arithmetic has no transcendental.
"""
import numpy as np
import pytest

from hyburn_amd import _lib as L
from hyburn_amd.engine import (BatchHybridEngine, MonitorConfig, ProjectionConfig, SimConfig,
                               SubnetworkConfig, SweepParam, SweepPoint, SweepSpec)
from hyburn_amd.instances import Instance, run_instances
from hyburn_amd.synth import DEFAULTS, demo_g2do_74, dense_to_csr, lognormal_connectome

pytestmark = pytest.mark.gpu


def _csr_net(N, seed, max_delay=12, dens=0.3):
    rng = np.random.default_rng(seed)
    W = lognormal_connectome(rng, N, dens, symmetric=False)
    ip, ix, w = dense_to_csr(W)
    d = rng.integers(1, max_delay + 1, len(ix)).astype(np.int32)
    return ip, ix, w, d


def _engine_cfg(model, N, x0, ip, ix, w, d, *, integ="heun", dt=0.1, T=200, cp=(0.05, 0.0), nsig=0.0,
                monitors=()):
    sub = SubnetworkConfig(model=model, nnodes=N, params=list(DEFAULTS[L.MODEL_NAMES[model]]),
                           initial_state=x0.ravel().tolist())
    proj = ProjectionConfig(src=0, tgt=0, weights={"indptr": ip.tolist(), "indices": ix.tolist(),
                                                   "data": w.tolist()},
                            delays=d.tolist(), coupling_fn="Linear", coupling_params=list(cp))
    return SimConfig(sim_length=T * dt, dt=dt, subnetworks=[sub], projections=[proj], integrator=integ,
                     nsig=nsig, monitors=list(monitors))


def _oracle_inst(model, N, x0, ip, ix, w, d, *, engine, delay_mode, T, dt=0.1, integ=L.HEUN,
                 cp=(0.05, 0.0), params=None, sweep_val=None, **kw):
    """The oracle instance of one engine sweep point: per-edge delays under the
    plain engine; the batch engine reads delays[0] through its sliding window
    whose horizon is max(delays) + 1 (engine.rs:253-270, 410, 541-589)."""
    m = L.MODEL_NAMES[model]
    args = dict(model=m, n_nodes=N, n_steps=T, dt=dt,
                params=np.array(DEFAULTS[m] if params is None else params, np.float32),
                x0=x0, integrator=integ, engine=engine, wkind=L.W_CSR, cfn=L.CF_LINEAR, cparams=cp,
                indptr=ip, indices=ix, weights=w, edge_delays=d, delay_mode=delay_mode,
                single_delay=int(d[0]), max_delay_hint=int(d.max()), record_traj=True)
    if sweep_val is not None:
        args.update(sweep_form=1, sweep_val=float(sweep_val))
    args.update(kw)
    return Instance(**args)


def test_run_sweep_batch_engine_bitwise(gpu, orc):
    """run_sweep_with_trajectory over the model's batch-native index (G2DO: I,
    dfun.rs:49-72) with the batch engine's delay read: [T][B][N][V] trajectory,
    tavg and final state bit-identical to the oracle."""
    N, T = 24, 300
    ip, ix, w, d = _csr_net(N, 5)
    x0 = np.random.default_rng(6).normal(0, 0.1, (N, 2)).astype(np.float32)
    eng = BatchHybridEngine.from_config(_engine_cfg("Generic2dOscillator", N, x0, ip, ix, w, d, T=T), n_sweep=6)
    vals = np.linspace(-0.5, 0.5, 6).astype(np.float32)
    r = eng.run_sweep_with_trajectory(SweepParam(0, 1), vals, T, exact=True)
    eng.close()
    traj = r.trajectories[0].reshape(T, 6, N, 2)
    tavg = r.tavg[0].reshape(6, N, 2)
    fin = r.final_states[0].reshape(6, N, 2)
    for b, v in enumerate(vals):
        I = _oracle_inst("Generic2dOscillator", N, x0, ip, ix, w, d, engine=L.ENGINE_BATCH,
                         delay_mode=L.DELAY_BATCH_COMPAT, T=T, sweep_val=v)
        o = orc.run(I)
        np.testing.assert_array_equal(traj[:, b], o["traj"], err_msg=f"point {b}")
        np.testing.assert_array_equal(tavg[b], o["tavg"])
        np.testing.assert_array_equal(fin[b], o["final_state"])


def test_run_sweep_non_native_index_is_ignored(gpu, orc):
    """A parameter index other than the model's batch-native one does not change
    the run (dfun_batch ignores it, dfun.rs:49-72)."""
    N, T = 16, 120
    ip, ix, w, d = _csr_net(N, 8)
    x0 = np.random.default_rng(9).normal(0, 0.1, (N, 2)).astype(np.float32)
    eng = BatchHybridEngine.from_config(_engine_cfg("Generic2dOscillator", N, x0, ip, ix, w, d, T=T), n_sweep=3)
    r = eng.run_sweep(SweepParam(0, 2), [-1.0, 0.0, 1.0], T, exact=True)
    eng.close()
    tavg = r.tavg[0].reshape(3, N, 2)
    o = orc.run(_oracle_inst("Generic2dOscillator", N, x0, ip, ix, w, d, engine=L.ENGINE_BATCH,
                             delay_mode=L.DELAY_BATCH_COMPAT, T=T))
    for b in range(3):
        np.testing.assert_array_equal(tavg[b], o["tavg"])


def test_run_sweep_plain_engine_per_edge_delays(gpu, orc):
    """The plain-engine mode (HybridEngine semantics) honours every edge's
    delay (runtime.rs:162-204)."""
    N, T = 20, 250
    ip, ix, w, d = _csr_net(N, 11, max_delay=30)
    x0 = np.random.default_rng(12).normal(0, 0.1, (N, 2)).astype(np.float32)
    cfg = _engine_cfg("MontbrioPazoRoxin", N, np.stack([np.full(N, 0.1), np.full(N, -2.0)], 1).astype(np.float32),
                      ip, ix, w, d, T=T, cp=(0.1, 0.0))
    eng = BatchHybridEngine.from_config(cfg, n_sweep=4, engine=L.ENGINE_PLAIN)
    vals = np.array([-5.0, -4.8, -4.6, -4.4], np.float32)
    r = eng.run_sweep_with_trajectory(SweepParam(0, 4), vals, T, exact=True)
    eng.close()
    traj = r.trajectories[0].reshape(T, 4, N, 2)
    x0m = np.stack([np.full(N, 0.1), np.full(N, -2.0)], 1).astype(np.float32)
    for b, v in enumerate(vals):
        I = _oracle_inst("MontbrioPazoRoxin", N, x0m, ip, ix, w, d, engine=L.ENGINE_PLAIN,
                         delay_mode=L.DELAY_PER_EDGE, T=T, cp=(0.1, 0.0), sweep_val=v)
        np.testing.assert_array_equal(traj[:, b], orc.run(I)["traj"], err_msg=f"point {b}")


def test_run_sweep_multi_cartesian_order(gpu, orc):
    """run_sweep_multi: Cartesian product, last dimension fastest
    (engine.rs:69-90); two model parameters take the per-point fallback
    (dfun.rs:99-133); a coupling parameter dimension sets G per point."""
    N, T = 18, 150
    ip, ix, w, d = _csr_net(N, 14)
    x0 = np.random.default_rng(15).normal(0, 0.1, (N, 2)).astype(np.float32)
    eng = BatchHybridEngine.from_config(_engine_cfg("Generic2dOscillator", N, x0, ip, ix, w, d, T=T), n_sweep=1)
    spec = SweepSpec([SweepPoint(0, 1, [0.0, 0.3]), SweepPoint(0, 2, [-1.5, 0.5]),
                      SweepPoint(0, 0, [0.01, 0.05, 0.1], kind=1)])
    r = eng.run_sweep_multi(spec, T, exact=True)
    eng.close()
    assert r.n_sweep == 12
    tavg = r.tavg[0].reshape(12, N, 2)
    s = 0
    for I_ in (0.0, 0.3):
        for a in (-1.5, 0.5):
            for G in (0.01, 0.05, 0.1):
                p = DEFAULTS[L.G2DO].copy()
                p[1], p[2] = I_, a
                inst = _oracle_inst("Generic2dOscillator", N, x0, ip, ix, w, d, engine=L.ENGINE_BATCH,
                                    delay_mode=L.DELAY_BATCH_COMPAT, T=T, params=p,
                                    cp=(np.float32(G), 0.0), record_traj=False)
                np.testing.assert_array_equal(tavg[s], orc.run(inst)["tavg"], err_msg=f"point {s}")
                s += 1


def test_run_sweep_bold_per_instance(gpu, orc):
    """A BOLD monitor on the engine gives one BOLD series per sweep point,
    [n_sweep][rows][N], resampled to TR (bold_monitor.rs:72-227).

    FP32 BOLD is coarse: v and q stay within a few ulps of 1 for the first
    windows, so the FP32 oracle's own error against the FP64 twin reaches
    ~1e-4 here and a one-ulp difference between device and host powf shows up
    at ~1e-6. The bar is therefore the FP64-relative one: the GPU's error is at
    most 4x the FP32 oracle's, + 1e-6; the neural states (tavg) are bit-exact."""
    N, T = 12, 4000
    ip, ix, w, d = _csr_net(N, 17)
    x0 = np.stack([np.full(N, 0.1), np.full(N, -2.0)], 1).astype(np.float32)
    mon = MonitorConfig(monitor_type="bold", tr=0.05, bold_period=10)
    cfg = _engine_cfg("MontbrioPazoRoxin", N, x0, ip, ix, w, d, T=T, cp=(0.1, 0.0), monitors=[mon])
    eng = BatchHybridEngine.from_config(cfg, n_sweep=3)
    vals = np.array([-5.0, -4.7, -4.4], np.float32)
    r = eng.run_sweep(SweepParam(0, 4), vals, T, exact=True)
    eng.close()
    bold = r.bold[0]
    tavg = r.tavg[0].reshape(3, N, 2)
    for b, v in enumerate(vals):
        I = _oracle_inst("MontbrioPazoRoxin", N, x0, ip, ix, w, d, engine=L.ENGINE_BATCH,
                         delay_mode=L.DELAY_BATCH_COMPAT, T=T, cp=(0.1, 0.0), sweep_val=v,
                         record_traj=False, bold_period=10, bold_tr_ms=50.0, record_bold=True)
        o32, o64 = orc.run(I), orc.run(I, dtype=np.float64)
        np.testing.assert_array_equal(tavg[b], o32["tavg"])
        assert bold.shape[1] == o32["bold"].shape[0] == I.bold_rows()
        e_gpu = np.max(np.abs(bold[b].astype(np.float64) - o64["bold"]))
        e_orc = np.max(np.abs(o32["bold"].astype(np.float64) - o64["bold"]))
        assert e_gpu <= 4 * e_orc + 1e-6, (b, e_gpu, e_orc)


def test_run_sweep_stochastic_shared_noise(gpu, orc):
    """Stochastic Heun with device Philox noise shared by all points (stream 0,
    the reference's one draw per step broadcast over the batch,
    engine.rs:729-741)."""
    N, T = 14, 200
    ip, ix, w, d = _csr_net(N, 19)
    x0 = np.random.default_rng(20).normal(0, 0.1, (N, 2)).astype(np.float32)
    cfg = _engine_cfg("Generic2dOscillator", N, x0, ip, ix, w, d, T=T, integ="heun_stochastic", nsig=0.01)
    eng = BatchHybridEngine.from_config(cfg, n_sweep=2, noise_seed=33)
    vals = np.array([0.1, 0.2], np.float32)
    r = eng.run_sweep_with_trajectory(SweepParam(0, 1), vals, T, exact=True)
    eng.close()
    traj = r.trajectories[0].reshape(T, 2, N, 2)
    for b, v in enumerate(vals):
        I = _oracle_inst("Generic2dOscillator", N, x0, ip, ix, w, d, engine=L.ENGINE_BATCH,
                         delay_mode=L.DELAY_BATCH_COMPAT, T=T, integ=L.HEUN_STOCH, sweep_val=v,
                         nsig=np.full(2, 0.01, np.float32), noise_seed=33, noise_stream=0)
        np.testing.assert_allclose(traj[:, b], orc.run(I)["traj"], rtol=1e-6, atol=1e-7)


def test_demo_g2do_74_plain_run_full_length(gpu, orc):
    """The plain forward workload at full length (G2DO, 74 nodes, 10k Heun steps,
    scalar self-coupling w*I): trajectory [10000][74][2] bit-identical to the
    oracle in exact mode; fast mode within 4x the FP32 oracle's own error
    against the FP64 twin."""
    I = demo_g2do_74()
    (g,), _ = run_instances([I], exact=True)
    o32 = orc.run(I)
    np.testing.assert_array_equal(g.traj, o32["traj"])
    np.testing.assert_array_equal(g.final_state, o32["final_state"])
    (gf,), _ = run_instances([I], exact=False)
    o64 = orc.run(I, dtype=np.float64)
    e_gpu = np.max(np.abs(gf.traj.astype(np.float64) - o64["traj"]))
    e_orc = np.max(np.abs(o32["traj"].astype(np.float64) - o64["traj"]))
    assert e_gpu <= 4 * e_orc + 1e-6, (e_gpu, e_orc)
