"""GPU tests on the benchmarked workloads themselves, through the C ABI.

* ragged_mpr_jr (4096 instances, MPR with BOLD / Jansen-Rit with legacy
  SigmoidalJansenRit, N in {68, 74, 84, 164, 400}, per-edge CSR delays, RK4
  at 0.05 ms): one instance of each of the ten (model, N) classes at 2 % of
  their simulated time, with BOLD, in one ragged plan -- against the oracle
  (exact mode: MPR trajectories bit-identical and BOLD within 4 FP32 ulps of
  the signal near rest, Jansen-Rit within 1e-5 over its first 200 steps; FMA
  mode within 4x the FP32 oracle's own error against FP64), plus
  the full-size plan's size-independent properties.
* fc_fit_mpr_84 (MPR 84 nodes, 10 % CSR, delays <= 30 ms, Heun 0.1 ms, BOLD,
  gradients in G and the 84 eta): 4 instances; gradients against FP64
  central differences of the oracle, and the checkpoint interval of the
  workload (500 steps) against the whole trajectory kept, at full length.
* kuramoto_order_1000 (Kuramoto 1000 nodes, 5 % CSR, delays <= 40 ms, RK4
  0.1 ms, order-parameter loss): 2 instances; gradients against FP64
  autograd of the restated graph and central differences, and sqrt(T)
  checkpointing against the whole trajectory kept, at full length.
"""
import dataclasses

import numpy as np
import pytest

from adjoint_cases import fc64, oracle_loss, rel_err
from hyburn_amd import _lib as L
from hyburn_amd.adjoint import AdjointPlan, Loss
from hyburn_amd.instances import run_instances
from hyburn_amd.synth import fc_fit_mpr_84, kuramoto_order_1000, ragged_mpr_jr

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _err_bound(g, o32, o64, atol=1e-6):
    e_gpu = np.max(np.abs(g.astype(np.float64) - o64))
    e_orc = np.max(np.abs(o32.astype(np.float64) - o64))
    assert e_gpu <= 4 * e_orc + atol, (e_gpu, e_orc)
    return e_gpu, e_orc


def _classes(t_scale=0.02):
    """The first instance of each (model, N) class of the 4096-instance batch."""
    insts = ragged_mpr_jr(n_inst=4096, t_scale=t_scale, bold=True)
    seen = {}
    for I in insts:
        seen.setdefault((I.model, I.n_nodes), I)
    # MPR instances run past one TR (720 ms = 14400 steps) so their BOLD is
    # resampled onto at least one row
    out = [dataclasses.replace(seen[k], record_traj=True,
                               n_steps=max(seen[k].n_steps, 15000) if k[0] == L.MPR else seen[k].n_steps)
           for k in sorted(seen)]
    assert len(out) == 10
    return out


@pytest.fixture(scope="module")
def classes():
    return _classes()


@pytest.fixture(scope="module")
def class_oracle(orc, classes):
    return [(orc.run(I), orc.run(I, dtype=np.float64)) for I in classes]


def test_ragged_classes_exact_mode(gpu, classes, class_oracle):
    print(f"[ragged classes exact] {[(I.model, I.n_nodes, I.n_steps) for I in classes]}", flush=True)
    outs, st = run_instances(classes, exact=True)
    assert st.n_launches >= 2
    for I, g, (o32, o64) in zip(classes, outs, class_oracle):
        tag = f"model={I.model} N={I.n_nodes} T={I.n_steps}"
        if I.model == L.MPR:
            np.testing.assert_array_equal(g.traj, o32["traj"], err_msg=tag)
            np.testing.assert_array_equal(g.tavg, o32["tavg"], err_msg=tag)
            np.testing.assert_array_equal(g.final_state, o32["final_state"], err_msg=tag)
            # BOLD runs through powf (device and host libm may differ in the last
            # ulp), and near rest the signal V0 (k1 (1 - q) + k2 (1 - q/v) + k3 (1 - v))
            # is a difference of O(1) terms: one ulp of q or v (1.2e-7) moves it by
            # V0 (k1 + k2) 1.2e-7 = 4.3e-6. Bar: 4 such ulps, and the FP64 error
            # bound of the FMA-mode tests
            assert g.bold.shape == o32["bold"].shape == (I.bold_rows(), I.n_nodes), tag
            np.testing.assert_allclose(g.bold, o32["bold"], rtol=0, atol=4 * 4.3e-6, err_msg=tag)
            _err_bound(g.bold, o32["bold"], o64["bold"], atol=4.3e-6)
        else:
            # Jansen-Rit's sigmoids run through expf
            np.testing.assert_allclose(g.traj[:200], o32["traj"][:200], rtol=1e-5, atol=1e-6, err_msg=tag)


def test_ragged_classes_fast_mode_error_bound(gpu, classes, class_oracle):
    outs, _ = run_instances(classes, exact=False, autotune=True)
    for I, g, (o32, o64) in zip(classes, outs, class_oracle):
        eg, eo = _err_bound(g.traj, o32["traj"], o64["traj"])
        print(f"[fast] model={I.model} N={I.n_nodes}: gpu {eg:.3e} oracle32 {eo:.3e}", flush=True)
        _err_bound(g.final_state, o32["final_state"], o64["final_state"])
        if I.model == L.MPR:
            assert g.bold.shape[0] >= 1
            _err_bound(g.bold, o32["bold"], o64["bold"])


def test_ragged_exact_mode_independent_of_batching(gpu, classes):
    """Each instance's result does not depend on which other instances share
    its plan (exact mode keeps each row's summation order)."""
    together, _ = run_instances(classes, exact=True)
    for I, g in zip(classes, together):
        (alone,), _ = run_instances([I], exact=True)
        np.testing.assert_array_equal(alone.traj, g.traj)
        if I.bold_period:
            np.testing.assert_array_equal(alone.bold, g.bold)


def test_ragged_full_size_properties(gpu):
    """The 4096-instance workload at full length (3.6e10 node-steps): finite
    outputs, one BOLD row per TR for every MPR instance, bitwise-repeatable
    runs, and a sample of instances equal to their single-instance plans."""
    from hyburn_amd.instances import Plan
    insts = ragged_mpr_jr(n_inst=4096, bold=True)
    print(f"[ragged full] {len(insts)} instances, {sum(I.n_nodes * I.n_steps for I in insts):.3e} node-steps",
          flush=True)
    import copy
    with Plan(insts, exact=True) as P:
        P.run()
        a = copy.deepcopy(P.download())
        P.run()
        b = copy.deepcopy(P.download())
    for I, x, y in zip(insts, a, b):
        assert np.all(np.isfinite(x.tavg)) and np.all(np.isfinite(x.final_state))
        np.testing.assert_array_equal(x.final_state, y.final_state)
        if I.bold_period:
            assert x.bold.shape == (I.bold_rows(), I.n_nodes) and np.all(np.isfinite(x.bold))
            np.testing.assert_array_equal(x.bold, y.bold)
    rng = np.random.default_rng(0)
    for i in rng.choice(len(insts), 6, replace=False):
        (alone,), _ = run_instances([insts[i]], exact=True)
        np.testing.assert_array_equal(alone.final_state, a[i].final_state)
        np.testing.assert_array_equal(alone.tavg, a[i].tavg)


# ------------------------------------------------------------ adjoint ------

def _grads(insts, losses, checkpoint_every=0):
    with AdjointPlan(insts, losses, checkpoint_every=checkpoint_every) as P:
        loss = P.forward()
        g = P.backward()
    return loss, g


def _fd(orc, I, kind, target, nodes, h=1e-4):
    """Central FP64 differences of the oracle loss in G and in eta of `nodes`."""
    def with_G(v):
        cp = list(I.cparams)
        cp[0] = v
        return dataclasses.replace(I, cparams=tuple(cp))
    g0 = float(np.float32(I.cparams[0]))
    gp, gm = float(np.float32(g0 + h)), float(np.float32(g0 - h))
    dG = (oracle_loss(orc, with_G(gp), kind, target) - oracle_loss(orc, with_G(gm), kind, target)) / (gp - gm)
    dP = []
    base = np.asarray(I.node_param, np.float32)
    for k in nodes:
        hk = h * max(1.0, abs(float(base[k])))
        up, dn = base.copy(), base.copy()
        up[k], dn[k] = np.float32(base[k] + hk), np.float32(base[k] - hk)
        lu = oracle_loss(orc, dataclasses.replace(I, node_param=up), kind, target)
        ld = oracle_loss(orc, dataclasses.replace(I, node_param=dn), kind, target)
        dP.append((lu - ld) / (float(up[k]) - float(dn[k])))
    return dG, np.array(dP)


def test_fc_fit_mpr_84_gradients_vs_fd(gpu, orc):
    """The workload's network and integrator at 2000 steps (200 ms, 200 native
    BOLD samples) with 4 instances in one plan. At this horizon every BOLD
    series still rides the common haemodynamic onset: the workload's FC sits
    at 1 - O(1e-6), and both ||FC - FC_target||^2 and the BOLD distance to the
    hidden-parameter run are ~1e-12, below FP32 resolution of the neural
    states. So the loss here is the BOLD MSE against 0.8 x each instance's own
    FP64 BOLD (residuals on the scale of the signal); the FC loss's adjoint is
    checked on a 40 s run in test_gpu_adjoint.py. Bar: 1e-4 for the loss,
    dL/dG and all 84 dL/deta."""
    insts, _ = fc_fit_mpr_84(n_inst=4, n_steps=2000)
    targets = []
    for I in insts:
        o = orc.run(dataclasses.replace(I, record_bold=True, bold_tr_ms=0.0), dtype=np.float64, native_bold=True)
        targets.append((0.8 * o["bold"]).astype(np.float32))
    loss, g = _grads(insts, [Loss(L.LOSS_BOLD_MSE, t, 2) for t in targets])
    for i, I in enumerate(insts):
        target = targets[i]
        lo = oracle_loss(orc, I, L.LOSS_BOLD_MSE, target)
        assert abs(loss[i] - lo) <= TOL * abs(lo), (i, loss[i], lo)
        fG, fP = _fd(orc, I, L.LOSS_BOLD_MSE, target, range(I.n_nodes))
        print(f"[fc_fit_mpr_84 #{i}] dG gpu {g[i].grad_coupling:.6e} fd {fG:.6e} | "
              f"deta rel {rel_err(g[i].grad_param, fP):.2e}", flush=True)
        assert rel_err(g[i].grad_coupling, fG) <= TOL
        assert rel_err(g[i].grad_param, fP) <= TOL


def test_fc_fit_mpr_84_checkpoint_every_500_full_length(gpu, orc):
    """Full length (20k steps) with the workload's FC loss: checkpoints every
    500 steps (segments recomputed in the reverse sweep) give the same loss
    and gradients, bit for bit, as the whole trajectory kept in HBM."""
    insts, hidden = fc_fit_mpr_84(n_inst=4)
    o = orc.run(dataclasses.replace(hidden, record_bold=True, bold_tr_ms=0.0), dtype=np.float64, native_bold=True)
    target = fc64(o["bold"].astype(np.float64)).astype(np.float32)
    losses = [Loss(L.LOSS_FC, target, 2)] * len(insts)
    la, ga = _grads(insts, losses, checkpoint_every=0)
    lb, gb = _grads(insts, losses, checkpoint_every=500)
    np.testing.assert_array_equal(la, lb)
    for x, y in zip(ga, gb):
        assert np.isfinite(x.grad_coupling) and np.all(np.isfinite(x.grad_param))
        assert x.grad_coupling == y.grad_coupling
        np.testing.assert_array_equal(x.grad_param, y.grad_param)
        np.testing.assert_array_equal(x.grad_x0, y.grad_x0)


def test_kuramoto_order_1000_gradients(gpu, orc):
    """The workload's network at 200 steps, 2 instances (G = 0 and 2): loss and
    gradients against FP64 autograd of the restated graph (G, the 1000
    natural frequencies, x0) and against FP64 central differences (G and
    eight frequencies)."""
    from oracle import hb_autodiff as ad
    insts, target = kuramoto_order_1000(n_inst=2, n_steps=200)
    tgt = target[:200]
    loss, g = _grads(insts, [Loss(L.LOSS_ORDER, tgt, 0)] * 2)
    for i, I in enumerate(insts):
        lt, gG, gP, gX = ad.gradients(I, L.LOSS_ORDER, tgt)
        print(f"[kuramoto_order_1000 #{i}] G={I.cparams[0]}: loss {loss[i]:.8e} vs {lt:.8e}; "
              f"dG {g[i].grad_coupling:.6e} vs {gG:.6e}; domega rel {rel_err(g[i].grad_param, gP):.2e}", flush=True)
        assert abs(loss[i] - lt) <= TOL * abs(lt)
        assert rel_err(g[i].grad_coupling, gG) <= TOL
        assert rel_err(g[i].grad_param, gP) <= TOL
        assert rel_err(g[i].grad_x0, gX) <= TOL
        nodes = list(range(0, 1000, 125))
        fG, fP = _fd(orc, I, L.LOSS_ORDER, tgt, nodes)
        assert rel_err(g[i].grad_coupling, fG) <= TOL
        assert rel_err(g[i].grad_param[nodes], fP) <= TOL


def test_kuramoto_order_1000_sqrt_checkpointing_full_length(gpu):
    """Full length (10k RK4 steps): sqrt(T) = 100-step checkpoints give the
    same loss and gradients, bit for bit, as the whole trajectory kept."""
    insts, target = kuramoto_order_1000(n_inst=2)
    losses = [Loss(L.LOSS_ORDER, target, 0)] * 2
    la, ga = _grads(insts, losses, checkpoint_every=0)
    lb, gb = _grads(insts, losses, checkpoint_every=100)
    np.testing.assert_array_equal(la, lb)
    for x, y in zip(ga, gb):
        assert np.isfinite(x.grad_coupling) and np.all(np.isfinite(x.grad_param))
        assert x.grad_coupling == y.grad_coupling
        np.testing.assert_array_equal(x.grad_param, y.grad_param)
