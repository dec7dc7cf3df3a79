"""GPU parity of the multi-subnetwork path (hb_run_multi, engine.rs:540-823)
against the oracle's restatement (oracle.hb_oracle.run_multi).

The kernel is built with FMA contraction and the device's expf/sinf, so it
is compared to the f32 oracle within rtol 2e-4, atol 2e-5 over short runs
(100-200 steps), and within rtol 2e-3 over the 1000-step ring workload.

FP64 bar: the oracle's FP64 twin (run_multi(..., dtype=float64), the same op
order in double) gives each result's error; the GPU's max abs error against
it must be at most 4x the f32 oracle's own, + 1e-6 (tavg, final states, BOLD).
"""
import numpy as np
import pytest

from hyburn_amd import _lib as L
from hyburn_amd.multi import Link, Subnet, run_multi
from hyburn_amd.synth import G2DO_DEFAULT, JR_DEFAULT, KUR_DEFAULT, MPR_DEFAULT, WC_DEFAULT

pytestmark = pytest.mark.gpu

RTOL, ATOL = 2e-4, 2e-5


def _x0(rng, n, nv, scale=0.3):
    return (rng.normal(size=(n, nv)) * scale).astype(np.float32)


def _bar(g, o32, o64, atol=1e-6):
    """GPU error against the FP64 twin <= 4x the f32 oracle's own + atol."""
    e_gpu = float(np.max(np.abs(np.asarray(g, np.float64) - o64))) if np.size(o64) else 0.0
    e_orc = float(np.max(np.abs(np.asarray(o32, np.float64) - o64))) if np.size(o64) else 0.0
    assert e_gpu <= 4 * e_orc + atol, (e_gpu, e_orc)
    return e_gpu, e_orc


def _check(res, tav, fin, rtol=RTOL, atol=ATOL, o64=None, bold=None):
    for k in range(len(tav)):
        assert np.all(np.isfinite(fin[k])) and np.all(np.isfinite(tav[k])), f"subnetwork {k} leaves f32 range"
        np.testing.assert_allclose(res.tavg[k], tav[k], rtol=rtol, atol=atol)
        np.testing.assert_allclose(res.final_states[k], fin[k], rtol=rtol, atol=atol)
    if o64 is not None:
        for k in range(len(tav)):
            _bar(res.tavg[k], tav[k], o64[0][k])
            _bar(res.final_states[k], fin[k], o64[1][k])
        if bold is not None:
            _bar(res.bold, bold, o64[2])


def _ring(n, w=0.01):
    subs = [(L.G2DO, n, G2DO_DEFAULT, np.zeros((n, 2), np.float32)),
            (L.JANSEN_RIT, n, JR_DEFAULT, np.zeros((n, 6), np.float32)),
            (L.WILSON_COWAN, n, WC_DEFAULT, np.zeros((n, 2), np.float32))]
    links = [(0, 1, w, L.CF_LINEAR, [1.0], [(0, 0)]), (1, 2, w, L.CF_LINEAR, [1.0], [(0, 0)]),
             (2, 0, w, L.CF_LINEAR, [1.0], [(0, 0)])]
    return subs, links


def _run(subs, links, **kw):
    return run_multi([Subnet(m, n, p, x0) for (m, n, p, x0) in subs],
                     [Link(s, t, w, cfn, cp, cm) for (s, t, w, cfn, cp, cm) in links], **kw)


def test_three_subnet_ring_hybrid_sweep(gpu, orc):
    # the G2DO -> JR -> WC ring of the reference's generic CUDA benchmark, at 8 nodes
    subs, links = _ring(8)
    vals = np.linspace(-0.5, 0.5, 3).astype(np.float32)
    res = _run(subs, links, dt=0.1, n_steps=200, integrator=L.HEUN, hybrid_integrator=True, n_sweep=3,
               sweep=(0, 1, vals))
    tav, fin = orc.run_multi(subs, links, 0.1, 200, L.HEUN, True, sweep=(0, 1, vals), n_sweep=3)
    o64 = orc.run_multi(subs, links, 0.1, 200, L.HEUN, True, sweep=(0, 1, vals), n_sweep=3, dtype=np.float64)
    _check(res, tav, fin, o64=o64)
    assert not np.allclose(res.final_states[0][0], res.final_states[0][2])   # the sweep acts


@pytest.mark.parametrize("integ,hybrid", [(L.EULER, False), (L.HEUN, False), (L.RK4, False), (L.RK4, True)])
def test_dense_links_and_cvar_remap(gpu, orc, integ, hybrid):
    rng = np.random.default_rng(integ * 2 + hybrid)
    subs = [(L.G2DO, 7, G2DO_DEFAULT, _x0(rng, 7, 2)), (L.MPR, 5, MPR_DEFAULT, np.abs(_x0(rng, 5, 2)))]
    W01 = (rng.uniform(size=(5, 7)) * 0.1).astype(np.float32)
    W10 = (rng.uniform(size=(7, 5)) * 0.1).astype(np.float32)
    links = [(0, 1, W01, L.CF_LINEAR, [1.0, 0.0], [(0, 1)]),        # G2DO V -> MPR coupling slot 1
             (1, 0, W10, L.CF_LINEAR, [0.5, 0.0], [(1, 0)]),        # MPR var 1 -> G2DO
             (1, 1, 0.05, L.CF_SCALED_LINEAR, [0.2, 0.1], [(0, 0)])]
    res = _run(subs, links, dt=0.05, n_steps=150, integrator=integ, hybrid_integrator=hybrid)
    tav, fin = orc.run_multi(subs, links, 0.05, 150, integ, hybrid)
    _check(res, tav, fin, o64=orc.run_multi(subs, links, 0.05, 150, integ, hybrid, dtype=np.float64))


def test_difference_kuramoto_sigmoidal(gpu, orc):
    rng = np.random.default_rng(7)
    subs = [(L.G2DO, 5, G2DO_DEFAULT, _x0(rng, 5, 2)), (L.G2DO, 7, G2DO_DEFAULT, _x0(rng, 7, 2)),
            (L.KURAMOTO, 6, KUR_DEFAULT, (rng.uniform(size=(6, 1)) * 6).astype(np.float32))]
    Wd = (rng.uniform(size=(7, 5)) * 0.2).astype(np.float32)
    Wk = (rng.uniform(size=(6, 6)) * 0.5).astype(np.float32)
    Ws = (rng.uniform(size=(5, 5)) * 0.2).astype(np.float32)
    links = [(0, 1, Wd, L.CF_DIFFERENCE, [0.3], [(0, 0)]),           # non-square: row sums
             (0, 0, Ws, L.CF_DIFFERENCE, [0.3], [(0, 0)]),           # square: diagonal rewrite
             (2, 2, Wk, L.CF_KURAMOTO, [1.5], [(0, 0)]),
             (1, 0, 0.1, L.CF_SIGMOIDAL, [], [(0, 0)]),
             (1, 1, 0.2, L.CF_DIFFERENCE, [0.4], [(0, 0)])]          # Scalar Difference: zero coupling
    res = _run(subs, links, dt=0.1, n_steps=120, integrator=L.HEUN)
    tav, fin = orc.run_multi(subs, links, 0.1, 120, L.HEUN, False)
    _check(res, tav, fin, o64=orc.run_multi(subs, links, 0.1, 120, L.HEUN, False, dtype=np.float64))


def test_ring_workload_full_length(gpu, orc):
    # 76 nodes per subnetwork, 1000 steps, 64 sweep points: point 0 and 63 against
    # the oracle, every point finite, tavg the mean of a bounded trajectory
    subs, links = _ring(76)
    vals = np.linspace(-0.5, 0.5, 64).astype(np.float32)
    res = _run(subs, links, dt=0.1, n_steps=1000, integrator=L.HEUN, hybrid_integrator=True, n_sweep=64,
               sweep=(0, 1, vals))
    for k in range(3):
        assert np.all(np.isfinite(res.tavg[k])) and np.all(np.isfinite(res.final_states[k]))
    for b in (0, 63):
        tav, fin = orc.run_multi(subs, links, 0.1, 1000, L.HEUN, True, sweep=(0, 1, vals[b:b + 1]), n_sweep=1)
        for k in range(3):
            np.testing.assert_allclose(res.tavg[k][b], tav[k][0], rtol=2e-3, atol=2e-4)
            np.testing.assert_allclose(res.final_states[k][b], fin[k][0], rtol=2e-3, atol=2e-4)
        if b == 0:
            t64, f64 = orc.run_multi(subs, links, 0.1, 1000, L.HEUN, True, sweep=(0, 1, vals[:1]), n_sweep=1,
                                     dtype=np.float64)
            for k in range(3):
                _bar(res.tavg[k][0], tav[k][0], t64[k][0])
                _bar(res.final_states[k][0], fin[k][0], f64[k][0])
    assert res.kernel_ms > 0.0


def test_engine_api_multi_config(gpu, orc):
    # the reference's benchmark config (examples/bench/hyburn_bench_generic_cuda.rs:118-135) at
    # 6 nodes, through BatchHybridEngine.from_config / run_sweep with hybrid_integrator
    from hyburn_amd.engine import BatchHybridEngine, SimConfig, SweepParam
    n = 6
    d = {"sim_length": 100.0, "dt": 0.1, "integrator": "Heun", "network": {
        "subnetworks": [
            {"model": "Generic2dOscillator", "nnodes": n, "params": G2DO_DEFAULT.tolist(), "initial_state": [0.0] * (2 * n)},
            {"model": "JansenRit", "nnodes": n, "params": JR_DEFAULT.tolist(), "initial_state": [0.0] * (6 * n)},
            {"model": "WilsonCowan", "nnodes": n, "params": WC_DEFAULT.tolist(), "initial_state": [0.0] * (2 * n)}],
        "projections": [{"src": a, "tgt": b, "weights": 0.01, "coupling_fn": "Linear", "coupling_params": [1.0],
                         "cvar_map": "0:0", "delays": [0]} for a, b in ((0, 1), (1, 2), (2, 0))]}}
    eng = BatchHybridEngine.from_config(SimConfig.from_dict(d), 4)
    eng.hybrid_integrator = True
    vals = np.linspace(-0.5, 0.5, 4).astype(np.float32)
    r = eng.run_sweep(SweepParam(sub_idx=0, param_idx=1), vals, 300)
    subs, links = _ring(n)
    tav, fin = orc.run_multi(subs, links, 0.1, 300, L.HEUN, True, sweep=(0, 1, vals), n_sweep=4)
    for k in range(3):
        np.testing.assert_allclose(r.tavg[k], tav[k].ravel(), rtol=RTOL, atol=ATOL)
        np.testing.assert_allclose(r.final_states[k], fin[k].ravel(), rtol=RTOL, atol=ATOL)


def test_delayed_projections(gpu, orc):
    # the ring with delays 5, 3, 7 steps plus a dense delayed link; the history
    # horizon comes from the largest delay of any list (here 11 > the links' own)
    rng = np.random.default_rng(11)
    n = 6
    subs = [(L.G2DO, n, G2DO_DEFAULT, _x0(rng, n, 2)), (L.JANSEN_RIT, n, JR_DEFAULT, _x0(rng, n, 6, 0.01)),
            (L.WILSON_COWAN, n, WC_DEFAULT, np.abs(_x0(rng, n, 2, 0.1)))]
    W = (rng.uniform(size=(n, n)) * 0.2).astype(np.float32)
    links = [(0, 1, 0.05, L.CF_LINEAR, [1.0], [(0, 0)], 5), (1, 2, 0.05, L.CF_LINEAR, [1.0], [(0, 0)], 3),
             (2, 0, 0.05, L.CF_LINEAR, [1.0], [(0, 0)], 7), (0, 0, W, L.CF_LINEAR, [0.3, 0.0], [(0, 0)], 2)]
    for md in (0, 11):
        res = run_multi([Subnet(m, k, p, x0) for (m, k, p, x0) in subs],
                        [Link(s, t, w, cfn, cp, cm, dl) for (s, t, w, cfn, cp, cm, dl) in links],
                        dt=0.1, n_steps=150, integrator=L.RK4, n_sweep=2, max_delay=md)
        tav, fin = orc.run_multi(subs, links, 0.1, 150, L.RK4, False, n_sweep=2, max_delay=md)
        _check(res, tav, fin, o64=orc.run_multi(subs, links, 0.1, 150, L.RK4, False, n_sweep=2, max_delay=md,
                                                dtype=np.float64))


@pytest.mark.parametrize("integ,hybrid", [(L.EULER_STOCH, False), (L.HEUN_STOCH, True), (L.RK4_STOCH, False)])
def test_stochastic_multi(gpu, orc, integ, hybrid):
    # one raw draw per node, variable and step, shared by the sweep points;
    # under hybrid_integrator G2DO steps with deterministic Heun
    rng = np.random.default_rng(integ)
    n, T = 6, 120
    subs = [(L.G2DO, n, G2DO_DEFAULT, _x0(rng, n, 2)), (L.WILSON_COWAN, n, WC_DEFAULT, np.abs(_x0(rng, n, 2, 0.1)))]
    links = [(0, 1, 0.05, L.CF_LINEAR, [1.0], [(0, 0)]), (1, 0, 0.05, L.CF_LINEAR, [1.0], [(0, 0)])]
    raw = [rng.standard_normal((T, n, 2)).astype(np.float32) for _ in subs]
    nsig = np.array([0.001, 0.002], np.float32)
    vals = np.array([0.0, 0.3], np.float32)
    res = run_multi([Subnet(m, k, p, x0) for (m, k, p, x0) in subs],
                    [Link(s, t, w, cfn, cp, cm) for (s, t, w, cfn, cp, cm) in links],
                    dt=0.1, n_steps=T, integrator=integ, hybrid_integrator=hybrid, n_sweep=2, sweep=(0, 1, vals),
                    nsig=nsig, noise=raw)
    tav, fin = orc.run_multi(subs, links, 0.1, T, integ, hybrid, sweep=(0, 1, vals), n_sweep=2, nsig=nsig, noise=raw)
    _check(res, tav, fin, o64=orc.run_multi(subs, links, 0.1, T, integ, hybrid, sweep=(0, 1, vals), n_sweep=2,
                                            nsig=nsig, noise=raw, dtype=np.float64))


def test_split_run_resumes_bit_for_bit(gpu):
    # 200 steps at once == 100 steps, then 100 more from the returned states,
    # step counter and history (delays cross the split; noise drawn per step)
    rng = np.random.default_rng(21)
    n = 5
    subs = [Subnet(L.G2DO, n, G2DO_DEFAULT, _x0(rng, n, 2)), Subnet(L.WILSON_COWAN, n, WC_DEFAULT,
                                                                   np.abs(_x0(rng, n, 2, 0.1)))]
    links = [Link(0, 1, 0.05, delay=7), Link(1, 0, 0.05, delay=3), Link(0, 0, 0.02, delay=12)]
    raw = [rng.standard_normal((200, n, 2)).astype(np.float32) for _ in subs]
    vals = np.array([0.0, 0.2, -0.2], np.float32)
    kw = dict(dt=0.1, integrator=L.HEUN_STOCH, n_sweep=3, sweep=(0, 1, vals), nsig=[0.001, 0.001])
    full = run_multi(subs, links, n_steps=200, noise=raw, **kw)
    a = run_multi(subs, links, n_steps=100, noise=[z[:100] for z in raw], **kw)
    b = run_multi(subs, links, n_steps=100, noise=[z[100:] for z in raw], step0=a.step_end,
                  state=a.final_states, hist=a.hist, **kw)
    for k in range(2):
        np.testing.assert_array_equal(b.final_states[k], full.final_states[k])
        np.testing.assert_allclose((a.tavg[k] + b.tavg[k]) * 0.5, full.tavg[k], rtol=1e-5, atol=1e-6)


def test_engine_checkpoint_resume(gpu, tmp_path):
    # BatchHybridEngine keeps states, step and history across run_sweep calls
    # (engine.rs:496-540); checkpoint() / resume() carry them through a file
    from hyburn_amd.engine import BatchHybridEngine, SimConfig, SweepParam
    n = 4
    d = {"sim_length": 10.0, "dt": 0.1, "integrator": "rk4", "network": {
        "subnetworks": [
            {"model": "Generic2dOscillator", "nnodes": n, "params": G2DO_DEFAULT.tolist(),
             "initial_state": (np.arange(2 * n) * 0.05).tolist()},
            {"model": "JansenRit", "nnodes": n, "params": JR_DEFAULT.tolist(), "initial_state": [0.0] * (6 * n)}],
        "projections": [{"src": 0, "tgt": 1, "weights": 0.02, "delays": [4]},
                        {"src": 1, "tgt": 0, "weights": 0.02, "delays": [2, 9]}]}}
    vals = np.array([0.0, 0.4], np.float32)
    one = BatchHybridEngine.from_config(SimConfig.from_dict(d), 2)
    r_full = one.run_sweep(SweepParam(0, 1), vals, 160)
    two = BatchHybridEngine.from_config(SimConfig.from_dict(d), 2)
    two.run_sweep(SweepParam(0, 1), vals, 90)
    path = str(tmp_path / "ck.bin")
    two.checkpoint(path)
    three = BatchHybridEngine.from_config(SimConfig.from_dict(d), 2)
    three.resume(path)
    r3 = three.run_sweep(SweepParam(0, 1), vals, 70)
    r2 = two.run_sweep(SweepParam(0, 1), vals, 70)
    for k in range(2):
        np.testing.assert_array_equal(r3.final_states[k], r_full.final_states[k])
        np.testing.assert_array_equal(r2.final_states[k], r_full.final_states[k])


@pytest.mark.parametrize("integ", [L.HEUN, L.RK4_STOCH])
def test_plain_engine_with_stimuli(gpu, orc, integ):
    # the plain HybridEngine's multi-subnetwork step: true-ring delays, Scalar
    # weights as w * I, the four stimulus patterns on coupling variable 0
    rng = np.random.default_rng(integ + 50)
    n, T = 5, 160
    subs = [(L.G2DO, n, G2DO_DEFAULT, _x0(rng, n, 2)), (L.JANSEN_RIT, n, JR_DEFAULT, _x0(rng, n, 6, 0.01)),
            (L.WILSON_COWAN, n, WC_DEFAULT, np.abs(_x0(rng, n, 2, 0.1)))]
    links = [(0, 1, 0.05, L.CF_LINEAR, [1.0], [(0, 0)], 4), (1, 2, 0.05, L.CF_LINEAR, [1.0], [(0, 0)], 2),
             (2, 0, 0.05, L.CF_LINEAR, [1.0], [(0, 0)], 6)]
    stims = [(0, "impulse", [2.0, 1.5]), (0, "sinusoid", [1.0, 0.2, 40.0, 0.1]), (1, "step", [3.0, 9.0, 0.5]),
             (2, "pulse_train", [0.5, 0.3, 2.0, 0.7]), (7, "step", [0.0, 5.0, 1.0])]   # target 7 is skipped
    raw = [rng.standard_normal((T, n, L.MODEL_INFO[m][0])).astype(np.float32) for (m, _, _, _) in subs]
    nsig = np.full(6, 0.0005, np.float32)
    res = run_multi([Subnet(m, k, p, x0) for (m, k, p, x0) in subs],
                    [Link(s, t, w, cfn, cp, cm, dl) for (s, t, w, cfn, cp, cm, dl) in links],
                    dt=0.1, n_steps=T, integrator=integ, plain=True, stimuli=stims, nsig=nsig, noise=raw)
    tav, fin = orc.run_multi(subs, links, 0.1, T, integ, False, plain=True, stimuli=stims, nsig=nsig, noise=raw)
    _check(res, tav, fin, o64=orc.run_multi(subs, links, 0.1, T, integ, False, plain=True, stimuli=stims, nsig=nsig,
                                            noise=raw, dtype=np.float64))


def test_trajectory_recording(gpu, orc):
    # run_sweep_with_trajectory's layout, per subnetwork [n_steps][n_sweep][n][nvar]
    # (engine.rs:779-783): row k is the state after step k
    subs, links = _ring(5)
    vals = np.array([-0.3, 0.3], np.float32)
    kw = dict(dt=0.1, integrator=L.HEUN, hybrid_integrator=True, n_sweep=2, sweep=(0, 1, vals))
    full = run_multi([Subnet(m, n, p, x0) for (m, n, p, x0) in subs],
                     [Link(s, t, w, cfn, cp, cm) for (s, t, w, cfn, cp, cm) in links], n_steps=120,
                     record_trajectory=True, **kw)
    part = run_multi([Subnet(m, n, p, x0) for (m, n, p, x0) in subs],
                     [Link(s, t, w, cfn, cp, cm) for (s, t, w, cfn, cp, cm) in links], n_steps=50, **kw)
    tav, fin = orc.run_multi(subs, links, 0.1, 50, L.HEUN, True, sweep=(0, 1, vals), n_sweep=2)
    for k in range(3):
        tr = full.trajectories[k]
        assert tr.shape == (120, 2, 5, L.MODEL_INFO[subs[k][0]][0])
        np.testing.assert_array_equal(tr[-1], full.final_states[k])
        np.testing.assert_array_equal(tr[49], part.final_states[k])
        np.testing.assert_allclose(tr[49], fin[k], rtol=RTOL, atol=ATOL)
        np.testing.assert_allclose(tr.mean(axis=0), full.tavg[k], rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("P,tr", [(10, 0.0), (7, 2.0), (25, 3.3)])
def test_bold_of_subnetwork_zero(gpu, orc, P, tr):
    # per sweep point, BOLD of subnetwork 0's variable 0 (the batch engine's
    # monitor target); windows, flush of a partial window, TR resampling
    rng = np.random.default_rng(P)
    subs, links = _ring(6)
    subs[0] = (L.G2DO, 6, G2DO_DEFAULT, _x0(rng, 6, 2))
    vals = np.array([-0.2, 0.4], np.float32)
    T = 233
    res = run_multi([Subnet(m, n, p, x0) for (m, n, p, x0) in subs],
                    [Link(s, t, w, cfn, cp, cm) for (s, t, w, cfn, cp, cm) in links], dt=0.1, n_steps=T,
                    integrator=L.HEUN, hybrid_integrator=True, n_sweep=2, sweep=(0, 1, vals), bold_period=P,
                    bold_tr_ms=tr)
    tav, fin, bold = orc.run_multi(subs, links, 0.1, T, L.HEUN, True, sweep=(0, 1, vals), n_sweep=2,
                                   bold_period=P, bold_tr_ms=tr)
    assert res.bold.shape == bold.shape and bold.shape[1] > 0
    np.testing.assert_allclose(res.bold, bold, rtol=RTOL, atol=ATOL)
    o64 = orc.run_multi(subs, links, 0.1, T, L.HEUN, True, sweep=(0, 1, vals), n_sweep=2, bold_period=P,
                        bold_tr_ms=tr, dtype=np.float64)
    _check(res, tav, fin, o64=o64, bold=bold)


def test_plain_per_edge_csr_delays(gpu, orc):
    # per-edge CSR delays between and within subnetworks (plain engine,
    # runtime.rs:125-205): each edge reads its source d_e steps back, zero
    # while d_e > step
    rng = np.random.default_rng(77)
    subs = [(L.G2DO, 6, G2DO_DEFAULT, _x0(rng, 6, 2)), (L.G2DO, 5, G2DO_DEFAULT, _x0(rng, 5, 2)),
            (L.WILSON_COWAN, 5, WC_DEFAULT, np.abs(_x0(rng, 5, 2, 0.1)))]

    def csr(nt, ns, md):
        W = (rng.uniform(size=(nt, ns)) * (rng.uniform(size=(nt, ns)) < 0.6)).astype(np.float32)
        ipt = np.concatenate([[0], np.cumsum((W != 0).sum(1))]).astype(np.int32)
        ind = np.concatenate([np.nonzero(W[r])[0] for r in range(nt)]).astype(np.int32)
        dat = np.concatenate([W[r][W[r] != 0] for r in range(nt)]).astype(np.float32)
        return ipt, ind, dat, rng.integers(0, md + 1, len(dat)).astype(np.int32)

    c00, c10, c21 = csr(6, 6, 9), csr(5, 6, 4), csr(5, 5, 13)
    links_o = [(0, 0, c00, L.CF_LINEAR, [0.3, 0.0], [(0, 0)]), (0, 1, c10, L.CF_SIGMOIDAL, [], [(0, 0)]),
               (1, 2, c21, L.CF_LINEAR, [0.2, 0.01], [(0, 0)]), (2, 0, 0.05, L.CF_LINEAR, [1.0], [(0, 0)], 3)]
    links = [Link(0, 0, c00[:3], L.CF_LINEAR, [0.3, 0.0], edge_delays=c00[3]),
             Link(0, 1, c10[:3], L.CF_SIGMOIDAL, [], edge_delays=c10[3]),
             Link(1, 2, c21[:3], L.CF_LINEAR, [0.2, 0.01], edge_delays=c21[3]),
             Link(2, 0, 0.05, L.CF_LINEAR, [1.0], delay=3)]
    # Scalar weights between equal-size subnetworks only in the plain engine: 2 -> 0 is 5 -> 6
    links_o[3] = (2, 0, (rng.uniform(size=(6, 5)) * 0.1).astype(np.float32), L.CF_LINEAR, [1.0], [(0, 0)], 3)
    links[3] = Link(2, 0, links_o[3][2], L.CF_LINEAR, [1.0], delay=3)
    res = run_multi([Subnet(m, k, p, x0) for (m, k, p, x0) in subs], links, dt=0.1, n_steps=120,
                    integrator=L.HEUN, plain=True)
    tav, fin = orc.run_multi(subs, links_o, 0.1, 120, L.HEUN, False, plain=True)
    _check(res, tav, fin, o64=orc.run_multi(subs, links_o, 0.1, 120, L.HEUN, False, plain=True, dtype=np.float64))


def test_linear_and_sup_hopf_subnetworks(gpu, orc):
    # Linear, SupHopf (dfun.rs:522-557), Hopfield and CoombesByrne2D run on the
    # multi-subnetwork path: a Hopf ring driving the others, with channel remaps
    from hyburn_amd.synth import LINEAR_DEFAULT, SUP_HOPF_DEFAULT
    rng = np.random.default_rng(88)
    subs = [(L.SUP_HOPF, 6, np.array([0.4, 2.0], np.float32), _x0(rng, 6, 2)),
            (L.G2DO, 6, G2DO_DEFAULT, _x0(rng, 6, 2)), (L.LINEAR, 6, LINEAR_DEFAULT, _x0(rng, 6, 1)),
            (L.HOPFIELD, 6, np.array([1.5, 3.0, 1.0], np.float32), _x0(rng, 6, 2))]
    W = (rng.uniform(size=(6, 6)) * 0.2).astype(np.float32)
    links = [(0, 0, W, L.CF_LINEAR, [0.3, 0.0], [(0, 0)]),            # 2 channels pass through
             (0, 1, 0.1, L.CF_LINEAR, [1.0], [(1, 0)]),               # Hopf y -> G2DO
             (1, 2, 0.2, L.CF_LINEAR, [1.0], [(0, 0)]),
             (2, 0, 0.05, L.CF_LINEAR, [1.0], [(0, 1)]),              # Linear x -> Hopf channel 1
             (0, 3, 0.3, L.CF_LINEAR, [1.0], [(0, 0), (1, 1)]),       # Hopf -> Hopfield, both channels
             (3, 3, W, L.CF_SIGMOIDAL, [], [(0, 0)])]
    for integ in (L.HEUN, L.RK4):
        res = _run(subs, links, dt=0.05, n_steps=150, integrator=integ, n_sweep=2)
        tav, fin = orc.run_multi(subs, links, 0.05, 150, integ, False, n_sweep=2)
        _check(res, tav, fin, o64=orc.run_multi(subs, links, 0.05, 150, integ, False, n_sweep=2, dtype=np.float64))
    assert SUP_HOPF_DEFAULT.shape == (2,)
    # CoombesByrne2D with a G2DO driver (r clamped at 0); v_syn = 0 keeps this
    # coupled network bounded (at the default v_syn = -4 it leaves f32 range
    # near step 100, in the oracle too)
    subs2 = [(L.COOMBES_BYRNE_2D, 6, np.array([1.0, 0.0, 1.0, 2.0], np.float32), np.abs(_x0(rng, 6, 2, 0.2))),
             (L.G2DO, 6, G2DO_DEFAULT, _x0(rng, 6, 2))]
    links2 = [(0, 0, 0.1, L.CF_LINEAR, [0.5, 0.0], [(0, 0)]),            # both channels pass through
              (1, 0, 0.05, L.CF_LINEAR, [1.0], [(0, 1)]), (0, 1, 0.05, L.CF_LINEAR, [1.0], [(0, 0)])]
    res = _run(subs2, links2, dt=0.02, n_steps=200, integrator=L.HEUN, n_sweep=2)
    tav, fin = orc.run_multi(subs2, links2, 0.02, 200, L.HEUN, False, n_sweep=2)
    assert all(np.all(np.isfinite(f)) for f in fin)
    _check(res, tav, fin, o64=orc.run_multi(subs2, links2, 0.02, 200, L.HEUN, False, n_sweep=2, dtype=np.float64))


# ---- the six fixture-pinned models (dfun.rs:952-1870) on the multi path ----

def _ext_cases():
    from hyburn_amd.synth import (EPI_CODIM3_DEFAULT, EPI_CODIM3_SLOWMOD_DEFAULT, EPI_RS_DEFAULT, KIONEX_DEFAULT,
                                  ZERLAUT_DEFAULT)
    return {
        "codim3": (L.EPILEPTOR_CODIM3, EPI_CODIM3_DEFAULT, [0.5, 0.0, 0.1], 0.1, L.HEUN, 300),
        "codim3_slowmod": (L.EPILEPTOR_CODIM3_SLOWMOD, EPI_CODIM3_SLOWMOD_DEFAULT, [0.5, 0.0, 0.05, 0.0, 0.0], 0.1,
                           L.RK4, 300),
        "epileptor_rs": (L.EPILEPTOR_RS, EPI_RS_DEFAULT, [-1.6, -12.5, 3.8, -1.0, 0.005, 0.0, 1.0, 0.0], 0.1, L.HEUN,
                         300),
        "zerlaut_first": (L.ZERLAUT_FIRST, ZERLAUT_DEFAULT, [0.01, 0.01, 0.0, 0.0, 0.0], 0.1, L.EULER, 200),
        "zerlaut_second": (L.ZERLAUT_SECOND, ZERLAUT_DEFAULT, [0.01, 0.01, 1e-4, 0.0, 1e-4, 0.0, 0.0, 0.0], 0.1,
                           L.EULER, 150),
        "kionex": (L.KIONEX, KIONEX_DEFAULT, [0.1, -65.0, 0.5, -0.5, -2.0], 0.01, L.HEUN, 600),
    }


@pytest.mark.parametrize("name", ["codim3", "codim3_slowmod", "epileptor_rs", "zerlaut_first", "zerlaut_second",
                                  "kionex"])
def test_fixture_models_on_the_gpu(gpu, orc, name):
    """Each model as a subnetwork of 6 nodes (spread initial states) driven
    by a G2DO subnetwork through a Linear link into coupling channel 0 and
    feeding back into it: GPU against the oracle (rtol 2e-4) and within 4x
    the f32 oracle's error against its FP64 twin."""
    m, p, base, dt, integ, T = _ext_cases()[name]
    rng = np.random.default_rng(len(name))
    nv = L.MODEL_INFO[m][0]
    x0 = (np.array(base, np.float32)[None, :] * (1.0 + 0.05 * rng.standard_normal((6, nv)))).astype(np.float32)
    subs = [(m, 6, p, x0), (L.G2DO, 5, G2DO_DEFAULT, _x0(rng, 5, 2))]
    links = [(1, 0, 0.02, L.CF_LINEAR, [1.0], [(0, 0)]), (0, 1, 0.01, L.CF_LINEAR, [1.0], [(0, 0)])]
    res = _run(subs, links, dt=dt, n_steps=T, integrator=integ, n_sweep=2)
    tav, fin = orc.run_multi(subs, links, dt, T, integ, False, n_sweep=2)
    _check(res, tav, fin, rtol=2e-4, atol=2e-5,
           o64=orc.run_multi(subs, links, dt, T, integ, False, n_sweep=2, dtype=np.float64))
    # the reference's trace test shape (two uncoupled nodes, plain engine): against the oracle, which the
    # CPU suite pins to the reference's TVB fixture rows
    two = [(m, 2, p, np.repeat(np.array(base, np.float32)[None, :], 2, 0))]
    r2 = run_multi([Subnet(*two[0])], [], dt=dt, n_steps=T, integrator=integ, plain=True)
    t2, f2 = orc.run_multi(two, [], dt, T, integ, False, plain=True)
    np.testing.assert_allclose(r2.final_states[0], f2[0], rtol=2e-4, atol=2e-5)


def test_fixture_models_hybrid_and_stochastic(gpu, orc):
    """hybrid_integrator: deterministic Heun for EpileptorRestingState, the
    Codim3 models and KIonEx (model_prefers_heun, dfun.rs:1871-1873), the
    configured stochastic Euler for Zerlaut; EpileptorRestingState takes its
    three coupling channels from a self-projection (channels pass through
    unmapped) and channel 0 from Codim3."""
    ext = _ext_cases()
    rng = np.random.default_rng(5)
    subs = []
    for name in ("epileptor_rs", "zerlaut_first", "codim3"):
        m, p, base, _, _, _ = ext[name]
        nv = L.MODEL_INFO[m][0]
        subs.append((m, 4, p, (np.array(base, np.float32)[None, :] *
                               (1.0 + 0.02 * rng.standard_normal((4, nv)))).astype(np.float32)))
    links = [(2, 0, 0.05, L.CF_LINEAR, [1.0], [(0, 0)]),              # codim3 x -> EpiRS channel 0
             (0, 0, 0.01, L.CF_LINEAR, [0.5, 0.0], [(0, 0)]),         # EpiRS x1, y1, z -> its 3 channels
             (0, 1, 0.001, L.CF_LINEAR, [1.0], [(0, 0)]), (1, 2, 0.1, L.CF_LINEAR, [1.0], [(0, 0)])]
    raw = [rng.standard_normal((120, 4, L.MODEL_INFO[s[0]][0])).astype(np.float32) for s in subs]
    nsig = np.full(8, 1e-6, np.float32)
    res = _run(subs, links, dt=0.05, n_steps=120, integrator=L.EULER_STOCH, hybrid_integrator=True, nsig=nsig,
               noise=raw)
    tav, fin = orc.run_multi(subs, links, 0.05, 120, L.EULER_STOCH, True, nsig=nsig, noise=raw)
    _check(res, tav, fin, o64=orc.run_multi(subs, links, 0.05, 120, L.EULER_STOCH, True, nsig=nsig, noise=raw,
                                            dtype=np.float64))


# ---- the remaining registry models (dfun.rs:609-950, 1212-1400) ----

def _golden_row0(name):
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "model_fixture_traces.npz")
    return np.load(path, allow_pickle=False)[name]


# model, the fixture whose first row is the base state, dt, steps, integrator
REMAINING = {
    "coombes_byrne": (L.COOMBES_BYRNE, "coombes_byrne_ref", 0.01, 100, L.RK4),
    "gast_sd": (L.GAST_SD, "gast_sd_ref", 0.1, 200, L.HEUN),
    "gast_sf": (L.GAST_SF, "gast_sf_ref", 0.01, 100, L.RK4),
    "larter_breakspear": (L.LARTER_BREAKSPEAR, "larter_breakspear_ref", 0.1, 200, L.EULER),
    "epileptor2d": (L.EPILEPTOR_2D, "epileptor2d_ref", 0.1, 200, L.HEUN),
    "epileptor": (L.EPILEPTOR, "epileptor_ref", 0.05, 200, L.RK4),
    "rww_exc_inh": (L.RWW_EXC_INH, "rww_exc_inh_ref", 0.1, 200, L.EULER),
    "deco_balanced": (L.DECO_BALANCED, "deco_balanced_ref", 0.1, 200, L.HEUN),
    "zetterberg_jansen": (L.ZETTERBERG_JANSEN, "zetterberg_jansen_ref", 0.1, 200, L.RK4),
    "reduced_fhn": (L.REDUCED_FHN, "reduced_fhn_ref", 0.1, 200, L.EULER),
    "reduced_hr": (L.REDUCED_HR, "reduced_hr_ref", 0.1, 200, L.HEUN),
    "dumont_gutkin": (L.DUMONT_GUTKIN, "dumont_gutkin_ref", 0.1, 200, L.RK4),
}


@pytest.mark.parametrize("name", list(REMAINING))
def test_remaining_models_on_the_gpu(gpu, orc, name):
    """Each model as a subnetwork of 6 nodes (the fixture's first row, spread)
    coupled both ways with a G2DO subnetwork through channel 0, 2 sweep
    points: GPU against the f32 oracle (rtol 2e-4) and within 4x the oracle's
    own error against its FP64 twin."""
    from hyburn_amd.synth import MULTI_DEFAULTS
    m, fx, dt, T, integ = REMAINING[name]
    nv = L.MODEL_INFO[m][0]
    rng = np.random.default_rng(len(name))
    base = _golden_row0(fx)[0]
    x0 = (base[None, :] * (1.0 + 0.05 * rng.standard_normal((6, nv)))).astype(np.float32)
    subs = [(m, 6, MULTI_DEFAULTS[m], x0), (L.G2DO, 5, G2DO_DEFAULT, _x0(rng, 5, 2))]
    links = [(1, 0, 0.02, L.CF_LINEAR, [1.0], [(0, 0)]), (0, 1, 0.01, L.CF_LINEAR, [1.0], [(0, 0)])]
    res = _run(subs, links, dt=dt, n_steps=T, integrator=integ, n_sweep=2)
    tav, fin = orc.run_multi(subs, links, dt, T, integ, False, n_sweep=2)
    _check(res, tav, fin, o64=orc.run_multi(subs, links, dt, T, integ, False, n_sweep=2, dtype=np.float64))
    assert not np.allclose(res.final_states[0][0], x0)   # the model moves


@pytest.mark.parametrize("name", ["coombes_byrne", "larter_breakspear", "epileptor2d", "epileptor", "rww_exc_inh",
                                  "deco_balanced", "reduced_fhn", "reduced_hr"])
def test_remaining_models_replay_their_fixture(gpu, name):
    """One uncoupled node from the fixture's first row, the plain engine's
    Euler at dt 0.1 with default parameters: the GPU follows the TVB trace the
    reference keeps (relative 5e-4 where |expected| >= 1e-4, absolute below),
    as the CPU suite checks for the oracle."""
    from hyburn_amd.synth import MULTI_DEFAULTS
    m, fx = REMAINING[name][0], REMAINING[name][1]
    ref = _golden_row0(fx)
    ok = np.isfinite(ref).all(1) & (np.abs(ref).max(1) < 1e4)
    n = int(np.argmin(ok)) if not ok.all() else ref.shape[0]
    r = run_multi([Subnet(m, 1, MULTI_DEFAULTS[m], ref[0][None, :])], [], dt=0.1, n_steps=n - 1,
                  integrator=L.EULER, plain=True, record_trajectory=True)
    got = r.trajectories[0][:, 0, 0, :].astype(np.float64)
    e = ref[1:n].astype(np.float64)
    big = np.abs(e) >= 1e-4
    assert np.all(np.isfinite(got))
    assert np.all(np.abs(got - e)[~big] < 1e-4)
    assert np.max(np.abs(got - e)[big] / np.abs(e)[big]) <= 5e-4


def test_four_cvar_and_twelve_variable_models_hybrid_stochastic(gpu, orc):
    """DumontGutkin's four coupling channels from a Linear self-projection
    (all four pass through), ZetterbergJansen's twelve variables (the
    twelve-register kernel), Epileptor's two channels from ReducedSetHindmarshRose
    (cvar_map 0:0 and 1:1); hybrid_integrator: deterministic Heun for
    Epileptor (model_prefers_heun), stochastic Euler elsewhere."""
    from hyburn_amd.synth import MULTI_DEFAULTS
    rng = np.random.default_rng(12)
    subs = []
    for name in ("dumont_gutkin", "zetterberg_jansen", "epileptor", "reduced_hr"):
        m, fx = REMAINING[name][0], REMAINING[name][1]
        nv = L.MODEL_INFO[m][0]
        base = _golden_row0(fx)[0]
        subs.append((m, 4, MULTI_DEFAULTS[m], (base[None, :] * (1.0 + 0.02 * rng.standard_normal((4, nv))))
                     .astype(np.float32)))
    links = [(0, 0, 0.02, L.CF_LINEAR, [0.5, 0.0], [(0, 0)]),   # DG: r_e, v_e, s_ee, s_ei -> its 4 channels
             (1, 0, 0.01, L.CF_LINEAR, [1.0], [(0, 0)]),        # ZJ v1 -> DG channel 0
             (3, 2, 0.01, L.CF_LINEAR, [1.0], [(0, 0), (1, 1)]),  # RHR xi, eta -> Epileptor channels 0, 1
             (2, 1, 0.001, L.CF_LINEAR, [1.0], [(0, 0)])]
    T = 120
    raw = [rng.standard_normal((T, 4, L.MODEL_INFO[s[0]][0])).astype(np.float32) for s in subs]
    nsig = np.full(12, 1e-6, np.float32)
    res = _run(subs, links, dt=0.05, n_steps=T, integrator=L.EULER_STOCH, hybrid_integrator=True, nsig=nsig,
               noise=raw, n_sweep=2)
    tav, fin = orc.run_multi(subs, links, 0.05, T, L.EULER_STOCH, True, nsig=nsig, noise=raw, n_sweep=2)
    _check(res, tav, fin, o64=orc.run_multi(subs, links, 0.05, T, L.EULER_STOCH, True, nsig=nsig, noise=raw,
                                            n_sweep=2, dtype=np.float64))


# ---- run_sweep_multi over several subnetworks (engine.rs:862-1348) ----

def test_cartesian_sweep_over_subnetworks(gpu, orc):
    subs, links = _ring(6)
    rng = np.random.default_rng(9)
    subs[0] = (L.G2DO, 6, G2DO_DEFAULT, _x0(rng, 6, 2))
    # G2DO I (native: swept form) x WC P (native) -> 6 points; then JR mu and J (two dims: per-point params)
    for spec, n, moved in (([(0, 1, [0.0, 0.25, -0.25]), (2, 18, [0.0, 0.5])], 6, 0),
                           ([(1, 12, [0.1, 0.3]), (1, 7, [120.0, 150.0]), (0, 2, [-1.0, -3.0])], 8, 1)):
        res = _run(subs, links, dt=0.1, n_steps=150, integrator=L.HEUN, hybrid_integrator=True, n_sweep=n,
                   sweep_spec=spec)
        tav, fin = orc.run_multi(subs, links, 0.1, 150, L.HEUN, True, n_sweep=n, sweep_spec=spec)
        _check(res, tav, fin, o64=orc.run_multi(subs, links, 0.1, 150, L.HEUN, True, n_sweep=n, sweep_spec=spec,
                                                dtype=np.float64))
        assert not np.allclose(res.final_states[moved][0], res.final_states[moved][-1])   # the sweep acts


def test_engine_run_sweep_multi_several_subnetworks(gpu, orc):
    """BatchHybridEngine.run_sweep_multi on a three-subnetwork config: every
    point's values in param_values, points from sweep point 0 of the current
    state, and the engine keeps the new states (a later run with the old
    n_sweep is an InvalidState)."""
    from hyburn_amd.engine import BatchHybridEngine, SimConfig, SweepParam, SweepPoint, SweepSpec
    n = 5
    d = {"sim_length": 10.0, "dt": 0.1, "integrator": "heun", "network": {
        "subnetworks": [
            {"model": "Generic2dOscillator", "nnodes": n, "params": G2DO_DEFAULT.tolist(),
             "initial_state": (np.arange(2 * n) * 0.03).tolist()},
            {"model": "JansenRit", "nnodes": n, "params": JR_DEFAULT.tolist(), "initial_state": [0.0] * (6 * n)},
            {"model": "WilsonCowan", "nnodes": n, "params": WC_DEFAULT.tolist(), "initial_state": [0.1] * (2 * n)}],
        "projections": [{"src": a, "tgt": b, "weights": 0.01, "coupling_fn": "Linear", "coupling_params": [1.0],
                         "cvar_map": "0:0", "delays": [0]} for a, b in ((0, 1), (1, 2), (2, 0))]}}
    eng = BatchHybridEngine.from_config(SimConfig.from_dict(d), 2)
    eng.run_sweep(SweepParam(0, 1), np.array([0.0, 0.3], np.float32), 40)
    start = [np.asarray(s).reshape(2, n, -1)[0] for s in eng._state]
    spec = SweepSpec([SweepPoint(0, 1, [0.1, 0.2, 0.3]), SweepPoint(2, 18, [0.0, 1.0])])
    r = eng.run_sweep_multi(spec, 60)
    assert r.n_sweep == 6 and r.param_values == [0.1, 0.0, 0.1, 1.0, 0.2, 0.0, 0.2, 1.0, 0.3, 0.0, 0.3, 1.0]
    subs = [(L.G2DO, n, G2DO_DEFAULT, start[0]), (L.JANSEN_RIT, n, JR_DEFAULT, start[1]),
            (L.WILSON_COWAN, n, WC_DEFAULT, start[2])]
    links = [(a, b, 0.01, L.CF_LINEAR, [1.0], [(0, 0)]) for a, b in ((0, 1), (1, 2), (2, 0))]
    tav, fin = orc.run_multi(subs, links, 0.1, 60, L.HEUN, False, n_sweep=6,
                             sweep_spec=[(0, 1, [0.1, 0.2, 0.3]), (2, 18, [0.0, 1.0])])
    for k in range(3):
        np.testing.assert_allclose(r.final_states[k], fin[k].ravel(), rtol=RTOL, atol=ATOL)
    assert eng._step == 60 and len(eng._state[0]) == 6
    with pytest.raises(L.HyburnError, match="sweep points"):
        eng.run_sweep(SweepParam(0, 1), np.array([0.0, 0.3], np.float32), 10)


def test_engine_batch_mean_bold_across_runs(gpu, orc):
    """The batch engine's BOLD monitor on the multi-subnetwork path
    (engine.rs:785-801): var0 of subnetwork 0 averaged over the sweep points in
    point order feeds one Balloon-Windkessel model per node, carried across two
    runs of 2005 and 1995 steps (partial windows flushed at each end), the
    second returning the whole series resampled to TR. A run_sweep_multi then
    leaves the engine without monitors (its local engine has none,
    engine.rs:899-905, 1331)."""
    from test_gpu_engine_state import BOLD_ULP, _monitor
    from hyburn_amd.engine import BatchHybridEngine, SimConfig, SweepParam, SweepPoint, SweepSpec
    n, P, dt = 6, 10, 0.1
    d = {"sim_length": 400.0, "dt": dt, "integrator": "Heun", "network": {
        "subnetworks": [
            {"model": "Generic2dOscillator", "nnodes": n, "params": G2DO_DEFAULT.tolist(),
             "initial_state": [0.1 * k for k in range(2 * n)]},
            {"model": "JansenRit", "nnodes": n, "params": JR_DEFAULT.tolist(), "initial_state": [0.0] * (6 * n)},
            {"model": "WilsonCowan", "nnodes": n, "params": WC_DEFAULT.tolist(), "initial_state": [0.0] * (2 * n)}],
        "projections": [{"src": a, "tgt": b, "weights": 0.01, "coupling_fn": "Linear", "coupling_params": [1.0],
                         "cvar_map": "0:0", "delays": [0]} for a, b in ((0, 1), (1, 2), (2, 0))]},
        "monitors": [{"monitor_type": "bold", "tr": 0.05, "bold_period": P}]}
    vals = np.array([-0.5, 0.1, 0.6], np.float32)
    eng = BatchHybridEngine.from_config(SimConfig.from_dict(d), 3, bold_batch_mean=True)
    a = eng.run_sweep_with_trajectory(SweepParam(sub_idx=0, param_idx=1), vals, 2005)
    b = eng.run_sweep_with_trajectory(SweepParam(sub_idx=0, param_idx=1), vals, 1995)

    def pmean(tr, T):   # var0 summed over the points in order, divided by their count (mean_dim)
        t = tr.reshape(T, 3, n, 2)[..., 0]
        s = t[:, 0].copy()
        for k in (1, 2):
            s = (s + t[:, k]).astype(np.float32)
        return (s / np.float32(3)).astype(np.float32)

    want = _monitor(orc, [pmean(a.trajectories[0], 2005), pmean(b.trajectories[0], 1995)], n, P, dt, 50.0)
    got = b.bold[0]
    # 201 + 200 native samples up to 400.5 ms: 8 rows at TR = 50 ms
    assert got.shape == (1, 8, n), got.shape
    np.testing.assert_allclose(got[0], want, rtol=0, atol=4 * BOLD_ULP)
    assert a.bold[0].shape == (1, 4, n)
    np.testing.assert_array_equal(a.bold[0], got[:, :4])
    eng.run_sweep_multi(SweepSpec([SweepPoint(0, 1, [-0.5, 0.0, 0.5])]), 50)
    c = eng.run_sweep(SweepParam(sub_idx=0, param_idx=1), vals, 50)
    assert c.bold is None
    eng.close()
