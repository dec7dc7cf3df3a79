"""`python -m hyburn_amd run`: the hyburn run command on the HIP backend
(src/cli.rs:785-842, 894-974, 1040-1103)."""
import os

import numpy as np
import pytest

from hyburn_amd import _lib as L
from hyburn_amd.cli import main, run
from hyburn_amd.instances import Instance

SIM = """
sim_length = 5.0
dt = 0.1
integrator = "heun"
[network]
[[network.subnetworks]]
model = "Generic2dOscillator"
nnodes = 6
params = [1, 0, -2, -10, 0, 0.02, 3, 1, 0, 1, 1, 1]
initial_state = [{ic}]
[[network.projections]]
src = 0
tgt = 0
weights = 0.01
coupling_fn = "Linear"
coupling_params = [0.01]
"""

SWEEP = """
parameter_name = "subnetworks[0].params[1]"
values = [0.0, 0.1, 0.2]
"""


def _files(tmp_path):
    ic = np.round(np.linspace(-0.3, 0.3, 12), 3)      # var-major [2][6] for the plain run
    cfg = tmp_path / "sim.toml"
    cfg.write_text(SIM.format(ic=", ".join(str(v) for v in ic)))
    sw = tmp_path / "sweep.toml"
    sw.write_text(SWEEP)
    return str(cfg), str(sw), ic.astype(np.float32)


def test_non_hip_backend_is_rejected(tmp_path):
    cfg, _, _ = _files(tmp_path)
    with pytest.raises(SystemExit) as e:
        run(cfg, str(tmp_path / "out"), backend="ndarray")
    assert "hip only" in str(e.value)


@pytest.mark.gpu
def test_plain_run_writes_reference_layout(gpu, orc, tmp_path):
    """state_traj.npy [T, nvar, nnodes, 1] and state_final_sub0.npy [nvar,
    nnodes, 1], var-major; values equal the plain-engine oracle run (scalar
    weights as w * I, construction.rs:639-655)."""
    cfg, _, ic = _files(tmp_path)
    out = tmp_path / "out"
    assert main(["run", "-c", cfg, "-o", str(out), "-b", "hip"]) == 0
    traj = np.load(out / "state_traj.npy")
    fin = np.load(out / "state_final_sub0.npy")
    assert traj.shape == (50, 2, 6, 1) and fin.shape == (2, 6, 1)
    x0 = ic.reshape(2, 6).T.copy()
    I = Instance(model=L.G2DO, n_nodes=6, n_steps=50, dt=0.1,
                 params=np.array([1, 0, -2, -10, 0, 0.02, 3, 1, 0, 1, 1, 1], np.float32), x0=x0,
                 integrator=L.HEUN, engine=L.ENGINE_PLAIN, wkind=L.W_SCALAR_DIAG, scalar_w=0.01,
                 cfn=L.CF_LINEAR, cparams=(0.01,), record_traj=True)
    o = orc.run(I)
    np.testing.assert_allclose(traj[:, :, :, 0], o["traj"].transpose(0, 2, 1), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(fin[:, :, 0], o["final_state"].T, rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_plain_run_checkpoint_then_resume(gpu, tmp_path):
    """cli.rs:1057-1071: --resume before the run, --checkpoint after it. 50
    steps, a checkpoint (the reference's version 2 file: one network, no
    delays), then 50 more from it equal the last 50 of a 100-step run."""
    from hyburn_amd import checkpoint as ck
    cfg, _, ic = _files(tmp_path)
    long_cfg = tmp_path / "sim100.toml"
    long_cfg.write_text(open(cfg).read().replace("sim_length = 5.0", "sim_length = 10.0"))
    ckpt = str(tmp_path / "a.ckpt")
    assert main(["run", "-c", cfg, "-o", str(tmp_path / "a"), "--checkpoint", ckpt]) == 0
    c = ck.read(ckpt)
    assert c["version"] == 2 and c["step"] == 50
    np.testing.assert_array_equal(c["states"][0][0].T, np.load(tmp_path / "a" / "state_final_sub0.npy")[:, :, 0])
    assert main(["run", "-c", cfg, "-o", str(tmp_path / "b"), "--resume", ckpt]) == 0
    assert main(["run", "-c", str(long_cfg), "-o", str(tmp_path / "w")]) == 0
    tb, tw = np.load(tmp_path / "b" / "state_traj.npy"), np.load(tmp_path / "w" / "state_traj.npy")
    assert tb.shape == (50, 2, 6, 1) and tw.shape == (100, 2, 6, 1)
    np.testing.assert_allclose(tb, tw[50:], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(np.load(tmp_path / "b" / "state_final_sub0.npy"),
                               np.load(tmp_path / "w" / "state_final_sub0.npy"), rtol=1e-6, atol=1e-7)


@pytest.mark.gpu
def test_plain_run_of_several_subnetworks_checkpoint_then_resume(gpu, tmp_path):
    """The multi-subnetwork plain run saves the batch file (version 3: states
    and history) and resumes from it; the stimulus clock runs on the global
    step, so 50 + 50 steps equal the last 50 of 100."""
    from hyburn_amd import checkpoint as ck
    cfg = tmp_path / "sim3.toml"
    cfg.write_text(SIM3)
    long_cfg = tmp_path / "sim3_100.toml"
    long_cfg.write_text(SIM3.replace("sim_length = 5.0", "sim_length = 10.0"))
    ckpt = str(tmp_path / "m.ckpt")
    assert main(["run", "-c", str(cfg), "-o", str(tmp_path / "a"), "--checkpoint", ckpt]) == 0
    assert ck.read(ckpt)["version"] == 3
    assert main(["run", "-c", str(cfg), "-o", str(tmp_path / "b"), "--resume", ckpt]) == 0
    assert main(["run", "-c", str(long_cfg), "-o", str(tmp_path / "w")]) == 0
    tb, tw = np.load(tmp_path / "b" / "state_traj.npy"), np.load(tmp_path / "w" / "state_traj.npy")
    np.testing.assert_allclose(tb, tw[tw.size // 2:], rtol=1e-6, atol=1e-7)
    for i in (0, 1):
        np.testing.assert_allclose(np.load(tmp_path / "b" / f"state_final_sub{i}.npy"),
                                   np.load(tmp_path / "w" / f"state_final_sub{i}.npy"), rtol=1e-6, atol=1e-7)


@pytest.mark.gpu
def test_sweep_run_writes_tavg_with_reference_header(gpu, tmp_path):
    """tavg_sub0.npy carries the header [n_sweep, nvar, nnodes] over the
    node-major [n_sweep, nnodes, nvar] data, as cli.rs:958-970 writes it."""
    from hyburn_amd.engine import BatchHybridEngine, SimConfig, SweepParam
    cfg, sw, _ = _files(tmp_path)
    out = tmp_path / "sw"
    assert main(["run", "-c", cfg, "-o", str(out), "--sweep", sw]) == 0
    raw = open(out / "tavg_sub0.npy", "rb").read()
    assert b"'shape': (3, 2, 6)" in raw
    t = np.load(out / "tavg_sub0.npy")
    eng = BatchHybridEngine.from_config(SimConfig.from_file(cfg), 3)
    r = eng.run_sweep(SweepParam(0, 1), [0.0, 0.1, 0.2], 50)
    eng.close()
    np.testing.assert_array_equal(t.ravel(), r.tavg[0])


SIM2 = """
sim_length = 5.0
dt = 0.1
integrator = "heun"
[network]
[[network.subnetworks]]
model = "Generic2dOscillator"
nnodes = 4
params = [1, 0, -2, -10, 0, 0.02, 3, 1, 0, 1, 1, 1]
initial_state = [0, 0, 0, 0, 0, 0, 0, 0]
[[network.subnetworks]]
model = "WilsonCowan"
nnodes = 4
params = [12, 4, 13, 11, 10, 10, 1.2, 2.8, 1, 0, 1, 4, 1, 0, 1, 1, 1, 1, 0, 0, 1, 1]
initial_state = [0, 0, 0, 0, 0, 0, 0, 0]
[[network.projections]]
src = 0
tgt = 1
weights = 0.01
coupling_fn = "Linear"
coupling_params = [1.0]
[[network.projections]]
src = 1
tgt = 0
weights = 0.01
coupling_fn = "Linear"
coupling_params = [1.0]
"""


@pytest.mark.gpu
def test_sweep_run_writes_one_tavg_per_subnetwork(gpu, tmp_path):
    # a two-subnetwork config runs on the multi-subnetwork path; cli.rs:958-970
    # writes tavg_sub{i}.npy for every subnetwork (Scalar weights need equal
    # node counts, as the config validation requires)
    from hyburn_amd.engine import BatchHybridEngine, SimConfig, SweepParam
    cfg = tmp_path / "sim2.toml"
    cfg.write_text(SIM2)
    sw = tmp_path / "sweep.toml"
    sw.write_text(SWEEP)
    out = tmp_path / "sw2"
    assert main(["run", "-c", str(cfg), "-o", str(out), "--sweep", str(sw)]) == 0
    assert b"'shape': (3, 2, 4)" in open(out / "tavg_sub0.npy", "rb").read()
    assert b"'shape': (3, 2, 4)" in open(out / "tavg_sub1.npy", "rb").read()
    eng = BatchHybridEngine.from_config(SimConfig.from_file(str(cfg)), 3)
    r = eng.run_sweep(SweepParam(0, 1), [0.0, 0.1, 0.2], 50)
    np.testing.assert_array_equal(np.load(out / "tavg_sub1.npy").ravel(), r.tavg[1])


SIM3 = SIM2 + """
[[stimuli]]
target = 1
temporal = "step"
params = [1.0, 3.0, 0.5]
"""


@pytest.mark.gpu
def test_plain_run_of_several_subnetworks(gpu, tmp_path):
    # cli.rs:1040-1103: state_final_sub{i}.npy per subnetwork as [nvar, nnodes, 1];
    # state_traj.npy is the flat per-step concatenation of every subnetwork's state
    from hyburn_amd.multi import Link, Subnet, run_multi
    from hyburn_amd.synth import G2DO_DEFAULT, WC_DEFAULT
    cfg = tmp_path / "sim3.toml"
    cfg.write_text(SIM3)
    out = tmp_path / "plain3"
    assert main(["run", "-c", str(cfg), "-o", str(out)]) == 0
    f0, f1 = np.load(out / "state_final_sub0.npy"), np.load(out / "state_final_sub1.npy")
    assert f0.shape == (2, 4, 1) and f1.shape == (2, 4, 1)
    tr = np.load(out / "state_traj.npy")
    assert tr.shape == (50 * (8 + 8),)
    r = run_multi([Subnet(L.G2DO, 4, G2DO_DEFAULT), Subnet(L.WILSON_COWAN, 4, WC_DEFAULT)],
                  [Link(0, 1, 0.01), Link(1, 0, 0.01)], dt=0.1, n_steps=50, integrator=L.HEUN, plain=True,
                  stimuli=[(1, "step", [1.0, 3.0, 0.5])])
    np.testing.assert_array_equal(f1[:, :, 0], r.final_states[1][0].T)
    np.testing.assert_array_equal(tr[-16:-8], r.final_states[0][0].T.ravel())


LINEAR_SIM = """
sim_length = 1.0
dt = 0.1
integrator = "heun"
[network]
[[network.subnetworks]]
model = "Linear"
nnodes = 3
params = [-10.0]
initial_state = [0.1, 0.2, 0.3]
[[network.projections]]
src = 0
tgt = 0
weights = 0.5
coupling_fn = "Linear"
coupling_params = [0.01]
"""


def test_plain_run_of_multi_only_model_takes_plain_multi(tmp_path, monkeypatch):
    """A plain run (no --sweep) of one subnetwork whose model only the
    multi-subnetwork kernel integrates goes through run_multi(plain=True):
    w * I Scalar weights and the true delay ring of the plain engine. The
    kernel call is stubbed, so this runs on the CPU."""
    import types

    from hyburn_amd import multi
    seen = {}

    def fake_run_multi(subs, links, **kw):
        seen.update(kw)
        n = kw["n_steps"]
        fin = [np.zeros((1, S.n_nodes, L.MODEL_INFO[S.model][0]), np.float32) for S in subs]
        trj = [np.zeros((n, 1, S.n_nodes, L.MODEL_INFO[S.model][0]), np.float32) for S in subs]
        return types.SimpleNamespace(final_states=fin, trajectories=trj)

    monkeypatch.setattr(multi, "run_multi", fake_run_multi)
    cfg = tmp_path / "lin.toml"
    cfg.write_text(LINEAR_SIM)
    out = run(str(cfg), str(tmp_path / "out"))
    assert seen["plain"] is True and seen["n_steps"] == 10
    assert np.load(out["state_final_sub0"]).shape == (1, 3, 1)


def test_from_config_multi_branch_rejects_single_network_options():
    from hyburn_amd.engine import BatchHybridEngine, SimConfig
    cfg = SimConfig.from_toml_str(LINEAR_SIM)
    for kw in (dict(engine=L.ENGINE_PLAIN), dict(delay_mode=L.DELAY_SINGLE), dict(noise_shared=False)):
        with pytest.raises(L.HyburnError) as e:
            BatchHybridEngine.from_config(cfg, 2, **kw)
        assert e.value.code == L.HB_INVALID_CONFIG


EXT_SIM = """
sim_length = {T}
dt = {dt}
integrator = "{integ}"
[network]
[[network.subnetworks]]
model = "{model}"
nnodes = 2
params = [{params}]
initial_state = [{ic}]
"""


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["EpileptorCodim3", "ZerlautAdaptationSecondOrder", "KIonEx"])
def test_plain_run_of_a_fixture_model(gpu, orc, tmp_path, name):
    """`hyburn run` of the reference's trace-test configuration
    (tests/model_reference_test.rs:90-156: two uncoupled nodes, the plain
    engine): state_final_sub0.npy [nvar, 2, 1] against the oracle, which the
    CPU suite pins to the reference's fixture rows."""
    from test_oracle_models_ext import TRACE_CASES
    m, p, ic, dt, integ, T, _, _ = TRACE_CASES[name]
    cfg = tmp_path / "sim.toml"
    cfg.write_text(EXT_SIM.format(T=dt * T, dt=dt, integ={L.HEUN: "heun", L.EULER: "euler"}[integ], model=name,
                                  params=", ".join(repr(float(v)) for v in p), ic=", ".join(str(v) for v in ic)))
    out = tmp_path / "out"
    assert main(["run", "-c", str(cfg), "-o", str(out), "-b", "hip"]) == 0
    fin = np.load(out / "state_final_sub0.npy")
    nv = L.MODEL_INFO[m][0]
    assert fin.shape == (nv, 2, 1)
    x0 = np.array(ic, np.float32).reshape(nv, 2).T.copy()
    _, f = orc.run_multi([(m, 2, p, x0)], [], dt, T, integ, False, plain=True)
    np.testing.assert_allclose(fin[:, :, 0], f[0][0].T, rtol=2e-4, atol=2e-5)
