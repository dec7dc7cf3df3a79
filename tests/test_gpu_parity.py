"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.

Bars:
* exact mode (HB_FLAG_EXACT, no FMA contraction): bit-identical to the FP32
  oracle for models whose arithmetic is +,-,*,/ and floor only (G2DO, MPR,
  Kuramoto with Linear coupling) -- this pins every index, delay slot, ring,
  noise and integrator bookkeeping step;
* models with exp/sin/cos/pow (device libm vs host libm differ in the last
  ulp): |gpu - oracle_f32| <= 1e-5 relative over short horizons;
* fast mode (FMA contraction): the GPU's error against the FP64 oracle is at
  most 4x the FP32 oracle's own error against FP64, plus 1e-6.
"""
import numpy as np
import pytest

from hyburn_amd import _lib as L
from hyburn_amd.instances import Instance, Plan, philox_normal as gpu_philox, run_instances
from hyburn_amd.synth import DEFAULTS, sweep_g2do_74, dense_to_csr, lognormal_connectome, philox_normal

pytestmark = pytest.mark.gpu

EXACT_MODELS = [L.G2DO, L.MPR, L.KURAMOTO]


def _x0(model, N, rng):
    nv = L.MODEL_INFO[model][0]
    if model == L.MPR:
        return (np.stack([np.full(N, 0.1), np.full(N, -2.0)], 1) + rng.normal(0, 0.01, (N, 2))).astype(np.float32)
    if model in (L.WILSON_COWAN, L.RWW):
        return rng.uniform(0.05, 0.5, (N, nv)).astype(np.float32)
    if model == L.KURAMOTO:
        return rng.uniform(0, 6.0, (N, 1)).astype(np.float32)
    return rng.normal(0, 0.1, (N, nv)).astype(np.float32)


def _cfn(model):
    if model == L.JANSEN_RIT:
        return L.CF_SJR, (1.0,)
    return L.CF_LINEAR, (0.05, 0.0)


def net(model, N=10, T=300, integ=L.HEUN, max_delay=15, dens=0.35, seed=0, **kw):
    rng = np.random.default_rng(seed)
    W = lognormal_connectome(rng, N, dens, symmetric=False)
    ip, ix, w = dense_to_csr(W)
    d = rng.integers(0, max_delay + 1, len(ix)).astype(np.int32)
    cfn, cp = _cfn(model)
    nv = L.MODEL_INFO[model][0]
    args = dict(model=model, n_nodes=N, n_steps=T, dt=0.05, params=DEFAULTS[model].copy(), x0=_x0(model, N, rng),
                integrator=integ, wkind=L.W_CSR, cfn=cfn, cparams=cp, indptr=ip, indices=ix, weights=w,
                edge_delays=d, delay_mode=L.DELAY_PER_EDGE, record_traj=True)
    if integ >= L.EULER_STOCH:
        args["nsig"] = np.full(nv, 0.001, np.float32)
        args["noise"] = rng.normal(size=(T, N, nv)).astype(np.float32)
    args.update(kw)
    return Instance(**args)


def _assert_exact(gpu_out, orc_out, tag=""):
    np.testing.assert_array_equal(gpu_out.traj, orc_out["traj"], err_msg=f"traj {tag}")
    np.testing.assert_array_equal(gpu_out.tavg, orc_out["tavg"], err_msg=f"tavg {tag}")
    np.testing.assert_array_equal(gpu_out.final_state, orc_out["final_state"], err_msg=f"final {tag}")


def _err_bound(g, o32, o64, atol=1e-6):
    e_gpu = np.max(np.abs(g.astype(np.float64) - o64))
    e_orc = np.max(np.abs(o32.astype(np.float64) - o64))
    assert e_gpu <= 4 * e_orc + atol, (e_gpu, e_orc)
    return e_gpu, e_orc


@pytest.mark.parametrize("model", EXACT_MODELS)
@pytest.mark.parametrize("integ", [L.EULER, L.HEUN, L.RK4, L.HEUN_STOCH, L.RK4_STOCH])
def test_exact_mode_bitwise(gpu, orc, model, integ):
    I = net(model, integ=integ, seed=10 + model * 7 + integ)
    (g,), _ = run_instances([I], exact=True)
    _assert_exact(g, orc.run(I), f"model={model} integ={integ}")


@pytest.mark.parametrize("model", [L.JANSEN_RIT, L.WILSON_COWAN, L.RWW])
@pytest.mark.parametrize("integ", [L.HEUN, L.RK4])
def test_transcendental_models_exact_mode(gpu, orc, model, integ):
    I = net(model, integ=integ, T=150, seed=3 + model)
    (g,), _ = run_instances([I], exact=True)
    o = orc.run(I)
    np.testing.assert_allclose(g.traj, o["traj"], rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("model", [L.G2DO, L.MPR, L.JANSEN_RIT, L.WILSON_COWAN, L.KURAMOTO, L.RWW])
def test_fast_mode_error_vs_fp64(gpu, orc, model):
    I = net(model, integ=L.RK4, T=200, seed=20 + model)
    (g,), _ = run_instances([I], exact=False)
    o32, o64 = orc.run(I), orc.run(I, dtype=np.float64)
    _err_bound(g.traj, o32["traj"], o64["traj"])


def test_kuramoto_coupling_fn(gpu, orc):
    rng = np.random.default_rng(4)
    N = 16
    W = lognormal_connectome(rng, N, 0.4)
    ip, ix, w = dense_to_csr(W)
    d = rng.integers(0, 12, len(ix)).astype(np.int32)
    I = Instance(model=L.KURAMOTO, n_nodes=N, n_steps=300, dt=0.1, params=np.array([0.02], np.float32),
                 x0=rng.uniform(0, 6, (N, 1)).astype(np.float32), integrator=L.RK4, wkind=L.W_CSR,
                 cfn=L.CF_KURAMOTO, cparams=(1.5,), indptr=ip, indices=ix, weights=w, edge_delays=d,
                 record_traj=True)
    (g,), _ = run_instances([I], exact=True)
    o32, o64 = orc.run(I), orc.run(I, dtype=np.float64)
    np.testing.assert_allclose(g.traj, o32["traj"], rtol=1e-5, atol=1e-5)
    (gf,), _ = run_instances([I], exact=False)
    _err_bound(gf.traj, o32["traj"], o64["traj"], atol=1e-5)


@pytest.mark.parametrize("kind", ["dense", "scalar_mean", "scalar_diag", "single_delay", "batch_compat"])
def test_coupling_kinds_exact(gpu, orc, kind):
    rng = np.random.default_rng(8)
    N = 12
    W = lognormal_connectome(rng, N, 0.5)
    base = dict(model=L.G2DO, n_nodes=N, n_steps=250, dt=0.1, params=DEFAULTS[L.G2DO], integrator=L.HEUN,
                x0=rng.normal(0, 0.2, (N, 2)).astype(np.float32), cfn=L.CF_LINEAR, cparams=(0.3, 0.01),
                record_traj=True)
    if kind == "dense":
        I = Instance(wkind=L.W_DENSE, dense=W, **base)
    elif kind == "scalar_mean":
        I = Instance(wkind=L.W_SCALAR_MEAN, scalar_w=0.7, **base)
    elif kind == "scalar_diag":
        I = Instance(wkind=L.W_SCALAR_DIAG, scalar_w=0.7, engine=L.ENGINE_PLAIN, **base)
    elif kind == "single_delay":
        I = Instance(wkind=L.W_DENSE, dense=W, delay_mode=L.DELAY_SINGLE, single_delay=9, **base)
    else:
        ip, ix, w = dense_to_csr(W)
        I = Instance(wkind=L.W_CSR, indptr=ip, indices=ix, weights=w, delay_mode=L.DELAY_BATCH_COMPAT,
                     single_delay=5, max_delay_hint=11, **base)
    (g,), _ = run_instances([I], exact=True)
    _assert_exact(g, orc.run(I), kind)


def test_difference_coupling_dense(gpu, orc):
    rng = np.random.default_rng(9)
    N = 8
    W = lognormal_connectome(rng, N, 0.6)
    I = Instance(model=L.G2DO, n_nodes=N, n_steps=200, dt=0.1, params=DEFAULTS[L.G2DO], integrator=L.HEUN,
                 x0=rng.normal(0, 0.2, (N, 2)).astype(np.float32), wkind=L.W_DENSE, dense=W,
                 cfn=L.CF_DIFFERENCE, cparams=(0.2,), record_traj=True)
    (g,), _ = run_instances([I], exact=True)
    _assert_exact(g, orc.run(I), "difference")


def test_node_parameter(gpu, orc):
    rng = np.random.default_rng(12)
    I = net(L.MPR, N=14, T=200, integ=L.HEUN, seed=12, node_param_idx=2,
            node_param=rng.uniform(-6, -4, 14).astype(np.float32))
    (g,), _ = run_instances([I], exact=True)
    _assert_exact(g, orc.run(I), "node param")


def test_swept_form(gpu, orc):
    I = net(L.G2DO, T=200, seed=13, sweep_form=1, sweep_val=0.37)
    (g,), _ = run_instances([I], exact=True)
    _assert_exact(g, orc.run(I), "swept")


def test_multi_nodes_per_thread(gpu, orc):
    """N > block size: node state table in LDS, nodes strided over threads."""
    I = net(L.G2DO, N=150, T=120, dens=0.1, seed=14)
    (g,), st = run_instances([I], exact=True, block_size=64)
    _assert_exact(g, orc.run(I), "multi")


def test_global_ring(gpu, orc):
    """A ring larger than LDS lives in HBM: 100 nodes x 602 slots x 4 B."""
    I = net(L.G2DO, N=100, T=700, dens=0.08, max_delay=600, seed=15)
    (g,), st = run_instances([I], exact=True)
    assert st.lds_bytes == 0
    _assert_exact(g, orc.run(I), "global ring")


def test_ragged_mixed_launch(gpu, orc):
    rng = np.random.default_rng(16)
    insts = []
    for i in range(14):
        model = [L.G2DO, L.MPR, L.KURAMOTO, L.JANSEN_RIT, L.WILSON_COWAN, L.RWW][i % 6]
        N = int(rng.choice([5, 33, 64, 70, 130]))
        T = int(rng.integers(50, 400))
        insts.append(net(model, N=N, T=T, integ=[L.HEUN, L.RK4][i % 2], dens=0.2, seed=100 + i))
    outs, st = run_instances(insts, exact=True)
    for I, g in zip(insts, outs):
        o = orc.run(I)
        if I.model in EXACT_MODELS:
            _assert_exact(g, o, f"ragged model={I.model} N={I.n_nodes}")
        else:
            np.testing.assert_allclose(g.traj, o["traj"], rtol=1e-5, atol=1e-6)


def test_persistent_queue_matches(gpu, orc):
    insts = [net(L.G2DO, N=20 + i, T=100 + 10 * i, seed=200 + i) for i in range(10)]
    a, _ = run_instances(insts, exact=True)
    b, _ = run_instances(insts, exact=True, grid=2)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.traj, y.traj)


def test_graph_replay_deterministic(gpu):
    insts = [net(L.MPR, N=30, T=300, integ=L.HEUN_STOCH, seed=300 + i) for i in range(6)]
    with Plan(insts, exact=False) as P:
        outs = []
        for _ in range(3):
            P.run()
            outs.append([o.final_state.copy() for o in P.download()])
    for r in outs[1:]:
        for x, y in zip(outs[0], r):
            np.testing.assert_array_equal(x, y)


def test_bold_per_instance(gpu, orc):
    for tr in (0.0, 2.5):   # native grid, then TR resampling
        I = net(L.MPR, N=12, T=1003, integ=L.HEUN, seed=17, bold_period=10, bold_tr_ms=tr, record_bold=True,
                record_traj=False)
        (g,), _ = run_instances([I], exact=True)
        o = orc.run(I)
        assert g.bold.shape == o["bold"].shape == (I.bold_rows(), 12)
        np.testing.assert_allclose(g.bold, o["bold"], rtol=1e-5, atol=1e-6)
        np.testing.assert_array_equal(g.tavg, o["tavg"])


def test_philox_gpu_matches_numpy(gpu):
    n = 1_000_003
    a = gpu_philox(77, 9, n)
    b = philox_normal(77, 9, n)
    assert np.mean(a == b) > 0.9999
    np.testing.assert_allclose(a, b, atol=1e-6)


def test_generated_noise_path(gpu, orc):
    """Noise generated on device from (seed, stream) equals the numpy stream the
    oracle is fed, so the trajectories agree."""
    I = net(L.G2DO, T=200, integ=L.HEUN_STOCH, seed=18, noise=None, noise_seed=5, noise_stream=3)
    (g,), _ = run_instances([I], exact=True)
    o = orc.run(I)
    np.testing.assert_allclose(g.traj, o["traj"], rtol=1e-6, atol=1e-7)


def test_sweep_g2do_74_full_length_exact_subset(gpu, orc):
    """The batched-sweep workload at full length (10k steps) for 8 of its 128
    points: bit-identical to the oracle in exact mode."""
    insts = sweep_g2do_74(n_sweep=128, explicit_noise=False)
    sub = [insts[i] for i in (0, 17, 40, 63, 64, 90, 111, 127)]
    for I in sub:
        I.noise = philox_normal(I.noise_seed, I.noise_stream, I.n_steps * I.n_nodes * 2).reshape(I.n_steps, I.n_nodes, 2)
    outs, _ = run_instances(sub, exact=True)
    for I, g in zip(sub, outs):
        o = orc.run(I)
        np.testing.assert_array_equal(g.tavg, o["tavg"])
        np.testing.assert_array_equal(g.final_state, o["final_state"])


# ---- delay-blocked schedule (per-edge delays with min delay >= 8) ----

def _blocked_net(model, N=24, T=400, integ=L.HEUN, dmin=8, dmax=60, dens=0.3, seed=0, **kw):
    rng = np.random.default_rng(seed)
    I = net(model, N=N, T=T, integ=integ, dens=dens, seed=seed, **kw)
    I.edge_delays = rng.integers(dmin, dmax + 1, len(I.indices)).astype(np.int32)
    return I


@pytest.mark.parametrize("model", EXACT_MODELS)
@pytest.mark.parametrize("integ", [L.EULER, L.HEUN, L.RK4, L.EULER_STOCH, L.HEUN_STOCH, L.RK4_STOCH])
def test_blocked_exact_bitwise(gpu, orc, model, integ):
    I = _blocked_net(model, integ=integ, seed=500 + 7 * model + integ)
    (g,), _ = run_instances([I], exact=True)
    _assert_exact(g, orc.run(I), f"blocked model={model} integ={integ}")


def test_blocked_matches_per_step_schedule(gpu):
    """The blocked and the one-barrier-per-step schedules give identical bits."""
    insts = [_blocked_net(m, integ=L.RK4, seed=600 + m) for m in (L.G2DO, L.MPR, L.JANSEN_RIT, L.RWW)]
    a, _ = run_instances(insts, exact=True)
    m = __import__("hyburn_amd.instances", fromlist=["_Marshal"])
    import ctypes as C
    mm = m._Marshal(insts)
    st = L.HbRunStats()
    o = m._opts(0, True, True, 0, 0)
    o.flags |= L.FLAG_NO_BLOCKING
    L.check(L.lib().hb_run_instances(mm.arr, len(insts), C.byref(o), C.byref(st)))
    for x, y in zip(a, mm.finish()):
        np.testing.assert_array_equal(x.traj, y.traj)


def test_blocked_global_ring_and_long_rows(gpu, orc):
    """Ring in HBM (300 nodes x 700 slots) and rows longer than 32 register edges."""
    I = _blocked_net(L.G2DO, N=300, T=300, dens=0.2, dmin=9, dmax=690, seed=601)
    assert I.indptr is not None and np.max(np.diff(I.indptr)) > 32
    (g,), st = run_instances([I], exact=True)
    assert st.lds_bytes == 0
    _assert_exact(g, orc.run(I), "blocked global ring")


@pytest.mark.parametrize("model", EXACT_MODELS)
def test_long_block_hbm_ring(gpu, orc, model):
    """Deterministic runs with the ring in HBM and every delay >= 16 use the
    16-step block (64-byte edge windows): bitwise equal to the oracle in exact
    mode, and within FP32 of it with four lanes per node."""
    I = _blocked_net(model, N=300, T=260, integ=L.RK4, dens=0.1, dmin=16, dmax=400, seed=610 + model)
    (g,), st = run_instances([I], exact=True)
    assert st.lds_bytes == 0
    o = orc.run(I)
    _assert_exact(g, o, f"long block model={model}")
    J = _blocked_net(model, N=120, T=260, integ=L.HEUN, dens=0.2, dmin=17, dmax=450, seed=620 + model)
    (f,), st = run_instances([J], exact=False)
    assert st.lds_bytes == 0
    np.testing.assert_allclose(f.traj, orc.run(J)["traj"], rtol=1e-5, atol=1e-5)
    # every delay >= 32: 32-step blocks (128-byte windows) for one-channel rings
    I = _blocked_net(model, N=300, T=260, integ=L.RK4, dens=0.1, dmin=32, dmax=400, seed=630 + model)
    (g,), st = run_instances([I], exact=True)
    assert st.lds_bytes == 0
    _assert_exact(g, orc.run(I), f"32-step block model={model}")


def test_blocked_jr_sjr_and_kuramoto_fn(gpu, orc):
    for I in (_blocked_net(L.JANSEN_RIT, integ=L.RK4, T=200, seed=602),
              _blocked_net(L.KURAMOTO, integ=L.RK4, T=300, seed=603, cfn=L.CF_KURAMOTO, cparams=(1.2,))):
        (g,), _ = run_instances([I], exact=True)
        np.testing.assert_allclose(g.traj, orc.run(I)["traj"], rtol=1e-5, atol=1e-5)


def test_long_block_jansen_rit(gpu, orc):
    """Jansen-Rit (6 variables, legacy sigmoidal coupling) with the ring in
    HBM on 16- and 32-step blocks, against the oracle."""
    for dmin in (16, 32):
        I = _blocked_net(L.JANSEN_RIT, N=300, T=200, integ=L.RK4, dens=0.1, dmin=dmin, dmax=400, seed=640 + dmin)
        (g,), st = run_instances([I], exact=True)
        assert st.lds_bytes == 0
        np.testing.assert_allclose(g.traj, orc.run(I)["traj"], rtol=1e-5, atol=1e-5)


def test_blocked_bold_and_node_params(gpu, orc):
    rng = np.random.default_rng(604)
    I = _blocked_net(L.MPR, N=20, T=1207, integ=L.HEUN, seed=604, bold_period=10, bold_tr_ms=2.5,
                     record_bold=True, record_traj=False, node_param_idx=2,
                     node_param=rng.uniform(-6, -4, 20).astype(np.float32))
    (g,), _ = run_instances([I], exact=True)
    o = orc.run(I)
    np.testing.assert_array_equal(g.tavg, o["tavg"])
    np.testing.assert_allclose(g.bold, o["bold"], rtol=1e-5, atol=1e-6)


def test_blocked_fast_mode_error(gpu, orc):
    I = _blocked_net(L.MPR, integ=L.RK4, T=300, seed=605)
    (g,), _ = run_instances([I], exact=False)
    _err_bound(g.traj, orc.run(I)["traj"], orc.run(I, dtype=np.float64)["traj"])


def test_blocked_lanes_per_node_fast_mode(gpu, orc):
    """FMA mode shares a node over up to 4 lanes; results stay within the FP32
    error bound and are reproducible run to run."""
    insts = [_blocked_net(L.G2DO, N=N, integ=L.HEUN_STOCH, T=300, seed=700 + N) for N in (7, 74, 130, 200)]
    a, _ = run_instances(insts, exact=False)
    b, _ = run_instances(insts, exact=False)
    for I, x, y in zip(insts, a, b):
        np.testing.assert_array_equal(x.traj, y.traj)
        _err_bound(x.traj, orc.run(I)["traj"], orc.run(I, dtype=np.float64)["traj"])


@pytest.mark.parametrize("dmin", [9, 20])
def test_blocked_two_channel_hbm_ring_bitwise(gpu, orc, dmin):
    """MPR with cv != 0 keeps both coupling channels; with the ring in HBM the
    channels are interleaved per slot (one contiguous window per edge). 8-step
    (dmin 9) and 16-step (dmin 20) blocks, bit-identical to the oracle."""
    p = DEFAULTS[L.MPR].copy()
    p[6] = 0.3
    I = _blocked_net(L.MPR, N=300, T=200, integ=L.RK4, dens=0.1, dmin=dmin, dmax=400, seed=650 + dmin, params=p)
    (g,), st = run_instances([I], exact=True)
    assert st.lds_bytes == 0
    _assert_exact(g, orc.run(I), f"two-channel HBM ring dmin={dmin}")


WIDE_CASES = {
    # (model, N, integrator, dmin, dmax, density): networks wider than 512
    # nodes on the 1024-thread delay-blocked kernel
    "kuramoto_linear_600_rk4_k8": (L.KURAMOTO, 600, L.RK4, 8, 40, 0.05),
    "g2do_700_heun_stoch_k8": (L.G2DO, 700, L.HEUN_STOCH, 9, 60, 0.04),
    "mpr_520_rk4_k16": (L.MPR, 520, L.RK4, 16, 200, 0.05),
    "mpr_530_heun_k32_as_k16": (L.MPR, 530, L.HEUN, 32, 300, 0.05),
}


@pytest.mark.parametrize("name", list(WIDE_CASES))
def test_wide_blocked_exact_bitwise(gpu, orc, name):
    """One node per thread over up to 1024 threads: bit-identical to the
    oracle in exact mode (LDS ring for the 8-step block, HBM ring for 16-step
    blocks; a 32-step choice becomes 16 steps on the wide kernel)."""
    model, N, integ, dmin, dmax, dens = WIDE_CASES[name]
    I = _blocked_net(model, N=N, T=90, integ=integ, dmin=dmin, dmax=dmax, dens=dens, seed=800 + N)
    (g,), st = run_instances([I], exact=True)
    assert st.block_size > 512, st.block_size
    _assert_exact(g, orc.run(I), name)


def test_wide_blocked_kuramoto_1000_fast_mode(gpu, orc):
    """The order-parameter workload's shape (Kuramoto coupling, 1000 nodes, 5%
    density, RK4, delays 9-400 steps) on the wide kernel in FMA mode: within
    the FP64 error bound, reproducible run to run."""
    rng = np.random.default_rng(811)
    N = 1000
    I = _blocked_net(L.KURAMOTO, N=N, T=60, integ=L.RK4, dmin=9, dmax=400, dens=0.05, seed=811,
                     cfn=L.CF_KURAMOTO, cparams=(1.5,), dt=0.1,
                     params=np.array([2 * np.pi * 10 / 1000], np.float32))
    I.x0 = rng.uniform(0, 2 * np.pi, (N, 1)).astype(np.float32)
    (a,), st = run_instances([I], exact=False)
    (b,), _ = run_instances([I], exact=False)
    assert st.block_size == 1024
    np.testing.assert_array_equal(a.traj, b.traj)
    _err_bound(a.traj, orc.run(I)["traj"], orc.run(I, dtype=np.float64)["traj"], atol=1e-5)


def test_autotuned_plan_matches_default(gpu, orc):
    """The launch-shape autotuner only changes lanes per node / block size:
    results stay within the FP32 bound of the default shapes, and every tuned
    class records one of its candidates."""
    from hyburn_amd.instances import autotune_entries
    from hyburn_amd.synth import ragged_mpr_jr
    L.lib().hb_autotune_clear()
    insts = ragged_mpr_jr(n_inst=24, t_scale=0.02, bold=False)
    for I in insts:
        I.record_traj = False
    a, _ = run_instances(insts, autotune=False)
    b, _ = run_instances(insts, autotune=True)
    ent = autotune_entries()
    assert len(ent) >= 1
    for e in ent:
        assert e["choice"] in {0: (256, 512), 1: (1, 2, 4), 2: (1, 16)}[e["blocked"]], e
    for I, x, y in zip(insts, a, b):
        np.testing.assert_allclose(y.final_state, x.final_state, rtol=1e-4, atol=1e-5)


# HyperbolicTangent and PreSigmoidal pre() run through tanhf: exact mode within
# 1e-5 of the FP32 oracle (device vs host libm), FMA mode within the FP64 bound
TANH_CASES = {
    "tanh_g2do_csr": dict(model=L.G2DO, cfn=L.CF_TANH, cparams=(0.02, 1.5, 0.1, 2.0)),
    "tanh_kuramoto_csr_rk4": dict(model=L.KURAMOTO, cfn=L.CF_TANH, cparams=(0.05, 0.5, 1.0, 1.0), integ=L.RK4),
    "presig_mpr_csr": dict(model=L.MPR, cfn=L.CF_PRESIGMOIDAL, cparams=(0.05, 0.5, 2.0, 0.8, 0.1)),
    "presig_g2do_csr_heun_stoch": dict(model=L.G2DO, cfn=L.CF_PRESIGMOIDAL, cparams=(0.03,), integ=L.HEUN_STOCH),
    "presig_mpr_cv_csr": dict(model=L.MPR, cfn=L.CF_PRESIGMOIDAL, cparams=(0.05, 0.5, 2.0, 0.8, 0.1),
                              params=np.array([1, 1, -5, 15, 0, 1, 0.3], np.float32)),
}


@pytest.mark.parametrize("name", list(TANH_CASES))
def test_tanh_and_presigmoidal_coupling(gpu, orc, name):
    kw = dict(TANH_CASES[name])
    model, integ = kw.pop("model"), kw.pop("integ", L.HEUN)
    I = net(model, integ=integ, T=250, seed=900 + len(name), **kw)
    (g,), _ = run_instances([I], exact=True)
    o32, o64 = orc.run(I), orc.run(I, dtype=np.float64)
    np.testing.assert_allclose(g.traj, o32["traj"], rtol=1e-5, atol=1e-5 if model == L.KURAMOTO else 1e-6)
    (gf,), _ = run_instances([I], exact=False)
    _err_bound(gf.traj, o32["traj"], o64["traj"], atol=1e-5 if model == L.KURAMOTO else 1e-6)


@pytest.mark.parametrize("wkind", [L.W_DENSE, L.W_SCALAR_MEAN, L.W_SCALAR_DIAG])
@pytest.mark.parametrize("cfn,cp", [(L.CF_TANH, (0.02, 1.5, 0.1, 2.0)), (L.CF_PRESIGMOIDAL, (0.05, 0.5, 2.0, 0.8, 0.1))])
def test_tanh_and_presigmoidal_weight_kinds(gpu, orc, wkind, cfn, cp):
    """Dense, mean-field and diagonal weights with one delay per projection,
    including the pre(0) fill of not-yet-simulated time (Q11)."""
    rng = np.random.default_rng(31 + wkind)
    N = 10
    W = lognormal_connectome(rng, N, 0.5)
    I = Instance(model=L.MPR, n_nodes=N, n_steps=200, dt=0.05, params=DEFAULTS[L.MPR].copy(),
                 x0=_x0(L.MPR, N, rng), integrator=L.HEUN, wkind=wkind, dense=W if wkind == L.W_DENSE else None,
                 scalar_w=0.3, cfn=cfn, cparams=cp, delay_mode=L.DELAY_SINGLE, single_delay=7, record_traj=True)
    (g,), _ = run_instances([I], exact=True)
    np.testing.assert_allclose(g.traj, orc.run(I)["traj"], rtol=1e-5, atol=1e-6)
