"""HYBURNCK checkpoints on the host (src/engine/runtime.rs:503-742).

* the reference's version 2 file is rebuilt here field by field from the
  layout runtime.rs:505-517 documents and 519-561 writes; this library's
  writer must produce the same bytes, and its reader must return the same
  values from them (the reference's state is [nvar][nnodes][nmodes],
  var-major; this library's is node-major);
* version 1 (scalar nsig) reads, as runtime.rs:640-646 reads it;
* the reference's error texts for a bad magic, a truncated file, an unknown
  version and an unknown integrator byte (runtime.rs:627-690);
* the single-network engine's state crosses the C boundary unchanged
  (hb_engine_get_state / hb_engine_set_state, host-only calls): a version 4
  file read into an engine and written back is the same file byte for byte;
  a reference file continues every sweep point from its one state and from
  a zero delay history; shape and history mismatches are refused.
The GPU suite checks that a run resumed from a file continues bit for bit.
"""
import struct

import numpy as np
import pytest

from hyburn_amd import _lib as L
from hyburn_amd import checkpoint as ck
from hyburn_amd.engine import BatchHybridEngine, ProjectionConfig, SimConfig, SubnetworkConfig
from hyburn_amd.synth import DEFAULTS


def _reference_v2(step, dt, ibyte, nsig, subs):
    """runtime.rs:505-517: subs = [(nvar, nnodes, nmodes, flat var-major values)]."""
    b = bytearray(b"HYBURNCK")
    b += (2).to_bytes(8, "little")
    b += int(step).to_bytes(8, "little")
    b += struct.pack("<d", dt)
    b += bytes([ibyte]) + bytes(7)
    b += len(nsig).to_bytes(4, "little") + bytes(4)
    for v in nsig:
        b += struct.pack("<f", v)
    b += len(subs).to_bytes(8, "little")
    for nv, nn, nm, _ in subs:
        b += nv.to_bytes(8, "little") + nn.to_bytes(8, "little") + nm.to_bytes(8, "little")
    for *_, flat in subs:
        for v in flat:
            b += struct.pack("<f", v)
    return bytes(b)


def test_reference_v2_bytes_and_read(tmp_path):
    rng = np.random.default_rng(1)
    nv, nn = 2, 5
    var_major = rng.normal(size=(nv, nn)).astype(np.float32)
    want = _reference_v2(1234, 0.05, 4, [0.01, 0.02], [(nv, nn, 1, var_major.ravel())])
    p = tmp_path / "a.ckpt"
    ck.write_reference(str(p), step=1234, dt=0.05, integrator=L.HEUN_STOCH, nsig=[0.01, 0.02], nvar=nv,
                       nnodes=nn, state=var_major.T)
    assert p.read_bytes() == want
    c = ck.read(str(p))
    assert c["version"] == 2 and c["step"] == 1234 and c["dt"] == 0.05 and c["integrator"] == L.HEUN_STOCH
    np.testing.assert_array_equal(c["nsig"], np.array([0.01, 0.02], np.float32))
    assert c["shapes"] == [(nv, nn, 1)] and c["n_sweep"] == 1
    np.testing.assert_array_equal(c["states"][0][0], var_major.T)
    assert c["hist"] is None and c["hist_single"] is None


def test_integrator_bytes_follow_the_reference():
    # runtime.rs:531-538
    assert ck.INTEGRATOR_BYTE == {L.HEUN: 1, L.EULER: 2, L.EULER_STOCH: 3, L.HEUN_STOCH: 4, L.RK4: 5,
                                  L.RK4_STOCH: 6}


def test_reference_v1_scalar_nsig(tmp_path):
    # v1: one f32 nsig and 4 padding bytes in place of the length and values
    b = bytearray(b"HYBURNCK") + (1).to_bytes(8, "little") + (7).to_bytes(8, "little") + struct.pack("<d", 0.1)
    b += bytes([2]) + bytes(7) + struct.pack("<f", 0.5) + bytes(4)
    b += (1).to_bytes(8, "little") + b"".join(v.to_bytes(8, "little") for v in (1, 3, 1))
    b += struct.pack("<3f", 1.0, 2.0, 3.0)
    p = tmp_path / "v1.ckpt"
    p.write_bytes(bytes(b))
    c = ck.read(str(p))
    assert c["step"] == 7 and c["integrator"] == L.EULER and list(c["nsig"]) == [0.5]
    np.testing.assert_array_equal(c["states"][0], np.array([[[1.0], [2.0], [3.0]]], np.float32))


@pytest.mark.parametrize("mutate, needle", [
    (lambda b: b"HYBURNCX" + b[8:], "bad magic"),
    (lambda b: b[:-3], "truncated"),
    (lambda b: b[:8] + (5).to_bytes(8, "little") + b[16:], "Unsupported checkpoint version: 5"),
    (lambda b: b[:32] + bytes([9]) + b[33:], "Unknown integrator kind in checkpoint: 9"),
])
def test_read_errors(tmp_path, mutate, needle):
    good = _reference_v2(3, 0.1, 1, [0.0], [(2, 2, 1, [1, 2, 3, 4])])
    p = tmp_path / "bad.ckpt"
    p.write_bytes(mutate(good))
    with pytest.raises(L.HyburnError) as e:
        ck.read(str(p))
    assert e.value.code == L.HB_INVALID_STATE and needle in str(e.value)


def _delayed_cfg(N=6, nsig=0.0, integ="heun"):
    rng = np.random.default_rng(7)
    W = (rng.random((N, N)) < 0.5).astype(np.float32) * rng.random((N, N)).astype(np.float32)
    ip = np.concatenate([[0], np.cumsum((W != 0).sum(1))]).astype(np.int32)
    ix = np.nonzero(W)[1].astype(np.int32)
    w = W[W != 0].astype(np.float32)
    d = rng.integers(1, 9, len(ix)).astype(np.int32)
    x0 = rng.normal(0, 0.3, (N, 2)).astype(np.float32)
    sub = SubnetworkConfig(model="Generic2dOscillator", nnodes=N, params=list(DEFAULTS[L.G2DO]),
                           initial_state=x0.ravel().tolist())
    proj = ProjectionConfig(src=0, tgt=0, weights={"indptr": ip.tolist(), "indices": ix.tolist(),
                                                   "data": w.tolist()},
                            delays=d.tolist(), coupling_fn="Linear", coupling_params=[0.05, 0.0])
    cfg = SimConfig(sim_length=10.0, dt=0.1, subnetworks=[sub], projections=[proj], integrator=integ, nsig=nsig)
    return cfg, x0, int(d.max())


def test_fresh_engine_writes_the_reference_file(tmp_path):
    """Before its first run an engine holds its configured state at step 0 and
    no history: the reference's own file says all of it."""
    cfg, x0, _ = _delayed_cfg()
    eng = BatchHybridEngine.from_config(cfg, 3)
    p = tmp_path / "fresh.ckpt"
    eng.checkpoint(str(p))
    c = ck.read(str(p))
    assert c["version"] == 2 and c["step"] == 0 and c["integrator"] == L.HEUN
    np.testing.assert_array_equal(c["states"][0][0], x0)
    eng.close()


def test_v4_file_round_trips_through_the_engine(tmp_path):
    cfg, x0, dmax = _delayed_cfg(nsig=0.003, integ="heun_stochastic")
    eng = BatchHybridEngine.from_config(cfg, 3)
    st = L.HbEngineState()
    L.check(L.lib().hb_engine_get_state(eng._h, st, None, None, None))
    N = 6
    assert (st.n_state, st.n_nodes, st.nvar) == (0, N, 2)
    # the last max(delays) steps of the one channel G2DO couples through
    assert st.hist_floats == dmax * N * 1
    rng = np.random.default_rng(3)
    states = rng.normal(size=(3, N, 2)).astype(np.float32)
    hist = rng.normal(size=(3, st.hist_floats)).astype(np.float32)
    a = tmp_path / "a.ckpt"
    ck.write_single(str(a), step=777, dt=0.05, integrator=L.RK4_STOCH, nsig=[0.01, 0.02], nvar=2, nnodes=N,
                    states=states, hist=hist)
    eng.resume(str(a))
    assert eng.step == 777 and eng.dt == 0.05
    b = tmp_path / "b.ckpt"
    eng.checkpoint(str(b))
    assert b.read_bytes() == a.read_bytes()
    # forced reference format: one point only
    with pytest.raises(L.HyburnError):
        eng.checkpoint(str(tmp_path / "c.ckpt"), reference_format=True)
    eng.close()


def test_reference_file_continues_every_point_from_a_zero_history(tmp_path):
    cfg, _, _ = _delayed_cfg()
    eng = BatchHybridEngine.from_config(cfg, 4)
    state = np.arange(12, dtype=np.float32).reshape(6, 2) / 10
    p = tmp_path / "ref.ckpt"
    p.write_bytes(_reference_v2(50, 0.1, 5, [0.0, 0.0], [(2, 6, 1, state.T.ravel())]))
    eng.resume(str(p))
    st, nsig, states, hist = eng._get_state()
    assert st.step == 50 and st.integrator == L.RK4 and st.n_state == 4
    np.testing.assert_array_equal(states, np.broadcast_to(state, (4, 6, 2)))
    assert hist.shape == (4, st.hist_floats) and not hist.any()
    # written back: four points and a history to carry, so the batch file
    q = tmp_path / "back.ckpt"
    eng.checkpoint(str(q))
    assert ck.read(str(q))["version"] == 4
    eng.close()


def test_resume_refuses_mismatches(tmp_path):
    cfg, _, _ = _delayed_cfg()
    eng = BatchHybridEngine.from_config(cfg, 2)
    p = tmp_path / "shape.ckpt"
    p.write_bytes(_reference_v2(1, 0.1, 1, [0.0], [(2, 5, 1, np.zeros(10))]))
    with pytest.raises(L.HyburnError) as e:
        eng.resume(str(p))
    assert "Checkpoint shape mismatch for subnetwork 0: expected (nvar=2,nnodes=6,nmodes=1), got (2,5,1)" in str(e.value)
    p.write_bytes(_reference_v2(1, 0.1, 1, [0.0], [(2, 6, 1, np.zeros(12))] * 2))
    with pytest.raises(L.HyburnError) as e:
        eng.resume(str(p))
    assert "Checkpoint has 2 subnetworks, expected 1" in str(e.value)
    q = tmp_path / "hist.ckpt"
    ck.write_single(str(q), step=1, dt=0.1, integrator=L.HEUN, nsig=[0.0], nvar=2, nnodes=6,
                    states=np.zeros((2, 6, 2), np.float32), hist=np.zeros((2, 5), np.float32))
    with pytest.raises(L.HyburnError) as e:
        eng.resume(str(q))
    assert "history holds 5 floats per point" in str(e.value)
    ck.write_single(str(q), step=1, dt=0.1, integrator=L.HEUN, nsig=[0.0], nvar=2, nnodes=6,
                    states=np.zeros((3, 6, 2), np.float32))
    with pytest.raises(L.HyburnError) as e:
        eng.resume(str(q))
    assert "holds 3 sweep points" in str(e.value)
    eng.close()
