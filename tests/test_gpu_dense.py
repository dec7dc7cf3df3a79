"""Dense all-to-all coupling on the matrix cores (hb_dense.hip) against the
oracle: batches of instances sharing one weight matrix, no delay, FMA mode.

The product W P^T runs on v_mfma_f32_16x16x4f32 in the matrix core's own
summation order, so the bar is the FP64-relative one of the FMA-mode tests:
the GPU's largest error against the FP64 twin is at most 4x the FP32
oracle's own, + atol. Both paths (matrix cores and, with HB_FLAG_NO_MFMA,
per-lane row sums) are checked; a batch of 24 instances fills one tile of 16
and pads a second, N = 37 pads the 16-target waves and the 4-wide k steps.
"""
import numpy as np
import pytest

from hyburn_amd import _lib as L
from hyburn_amd.instances import Instance, autotune_entries
from hyburn_amd.synth import DEFAULTS, lognormal_connectome

pytestmark = pytest.mark.gpu


def _x0(model, N, rng):
    nv = L.MODEL_INFO[model][0]
    if model == L.MPR:
        return (np.stack([np.full(N, 0.1), np.full(N, -2.0)], 1) + rng.normal(0, 0.01, (N, 2))).astype(np.float32)
    if model in (L.WILSON_COWAN, L.RWW):
        return rng.uniform(0.05, 0.5, (N, nv)).astype(np.float32)
    if model == L.KURAMOTO:
        return rng.uniform(0, 6.0, (N, 1)).astype(np.float32)
    return rng.normal(0, 0.1, (N, nv)).astype(np.float32)


def _batch(model, B=24, N=37, T=300, integ=L.HEUN, cfn=L.CF_LINEAR, cp=(0.05, 0.0), seed=0, dt=0.05):
    rng = np.random.default_rng(seed)
    W = lognormal_connectome(rng, N, 0.3, symmetric=False)
    nv = L.MODEL_INFO[model][0]
    sweep = {L.G2DO: (-0.2, 0.2), L.MPR: (-0.5, 0.5), L.JANSEN_RIT: (0.1, 0.3), L.WILSON_COWAN: (0.0, 0.5),
             L.KURAMOTO: (0.5, 1.5), L.RWW: (0.2, 0.4)}[model]
    vals = np.linspace(*sweep, B)
    out = []
    for b in range(B):
        kw = {}
        if integ >= L.EULER_STOCH:
            kw = dict(nsig=np.full(nv, 1e-3, np.float32), noise=rng.standard_normal((T, N, nv)).astype(np.float32))
        out.append(Instance(model=model, n_nodes=N, n_steps=T, dt=dt, params=DEFAULTS[model].copy(),
                            x0=_x0(model, N, rng), integrator=integ, wkind=L.W_DENSE, dense=W, cfn=cfn,
                            cparams=cp, sweep_form=1, sweep_val=float(vals[b]), record_traj=True, **kw))
    return out


def _run(insts, no_mfma=False):
    import ctypes as C

    from hyburn_amd.instances import _Marshal, _opts
    m = _Marshal(insts)
    o = _opts(0, False, True, 0, 0)
    if no_mfma:
        o.flags |= L.FLAG_NO_MFMA
    st = L.HbRunStats()
    L.check(L.lib().hb_run_instances(m.arr, len(insts), C.byref(o), C.byref(st)))
    return m.finish(), st


CASES = {
    "g2do_heun": (L.G2DO, L.HEUN, L.CF_LINEAR, (0.05, 0.0)),
    "g2do_rk4_stoch": (L.G2DO, L.RK4_STOCH, L.CF_LINEAR, (0.05, 0.0)),
    "mpr_heun": (L.MPR, L.HEUN, L.CF_LINEAR, (0.05, 0.0)),
    "jr_rk4_sjr": (L.JANSEN_RIT, L.RK4, L.CF_SJR, (1.0,)),
    "wc_heun_sigmoidal": (L.WILSON_COWAN, L.HEUN, L.CF_SIGMOIDAL, (0.0, 1.0, 0.0, 1.0, 1.0)),
    "kuramoto_rk4_kuramoto": (L.KURAMOTO, L.RK4, L.CF_KURAMOTO, (1.5,)),
    "rww_euler_stoch": (L.RWW, L.EULER_STOCH, L.CF_LINEAR, (0.3, 0.0)),
    "g2do_heun_difference": (L.G2DO, L.HEUN, L.CF_DIFFERENCE, (0.1,)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_dense_matrix_cores_vs_fp64(gpu, orc, name):
    model, integ, cfn, cp = CASES[name]
    insts = _batch(model, integ=integ, cfn=cfn, cp=cp, seed=len(name))
    g, st = _run(insts)
    assert st.block_size == 3 * 64, "the batch ran on the matrix-core kernel (3 waves of 16 targets)"
    gl, stl = _run(insts, no_mfma=True)
    assert stl.block_size == 64
    atol = 1e-5 if model == L.KURAMOTO else 1e-6
    for b in (0, 7, 15, 16, 23):   # both tiles, first and last slots
        I = insts[b]
        o32, o64 = orc.run(I), orc.run(I, dtype=np.float64)
        e_orc = np.max(np.abs(o32["traj"].astype(np.float64) - o64["traj"]))
        e_mfma = np.max(np.abs(g[b].traj.astype(np.float64) - o64["traj"]))
        e_lane = np.max(np.abs(gl[b].traj.astype(np.float64) - o64["traj"]))
        print(f"[dense {name} #{b}] mfma {e_mfma:.3e} lanes {e_lane:.3e} oracle32 {e_orc:.3e}", flush=True)
        assert e_mfma <= 4 * e_orc + atol, (b, e_mfma, e_orc)
        assert e_lane <= 4 * e_orc + atol, (b, e_lane, e_orc)
        np.testing.assert_allclose(g[b].tavg, o32["tavg"], rtol=1e-3, atol=1e-4)


def test_dense_small_batch_keeps_lane_loop_and_exact_mode(gpu, orc):
    """Fewer than 8 instances, or exact mode: the per-lane row sums (exact mode
    bit-identical to the FP32 oracle for G2DO)."""
    insts = _batch(L.G2DO, B=5, T=100, seed=3)
    g, st = _run(insts)
    assert st.block_size == 64
    from hyburn_amd.instances import run_instances
    ins = _batch(L.G2DO, B=20, T=100, seed=4)
    ge, _ = run_instances(ins, exact=True)
    for b in (0, 19):
        np.testing.assert_array_equal(ge[b].traj, orc.run(ins[b])["traj"])


def test_dense_autotuner_records_a_choice(gpu):
    L.lib().hb_autotune_clear()
    insts = _batch(L.G2DO, B=32, T=200, seed=5)
    from hyburn_amd.instances import run_instances
    run_instances(insts, autotune=True)
    ent = [e for e in autotune_entries() if e["blocked"] == 2]
    assert len(ent) == 1 and ent[0]["choice"] in (1, 16), ent
