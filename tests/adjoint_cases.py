"""Shared cases and checkers for the adjoint tests: seeded small networks,
losses computed from the FP64 oracle's outputs, and central finite
differences of those losses."""
from __future__ import annotations

import dataclasses

import numpy as np

from hyburn_amd import _lib as L
from hyburn_amd.instances import Instance
from hyburn_amd.synth import DEFAULTS, dense_to_csr, lognormal_connectome


def _net(rng, N, dens, dmin, dmax):
    W = lognormal_connectome(rng, N, dens, symmetric=False)
    ip, ix, w = dense_to_csr(W)
    d = rng.integers(dmin, dmax + 1, len(ix)).astype(np.int32)
    return ip, ix, w, d


def kuramoto_case(N=6, T=120, seed=1, integ=L.RK4, dmin=1, dmax=8, G=1.5, dens=0.5):
    rng = np.random.default_rng(seed)
    ip, ix, w, d = _net(rng, N, dens, dmin, dmax)
    omega = (2 * np.pi * rng.normal(10.0, 1.0, N) / 1000.0).astype(np.float32)
    I = Instance(model=L.KURAMOTO, n_nodes=N, n_steps=T, dt=0.1, params=np.array([0.0], np.float32),
                 x0=rng.uniform(0, 2 * np.pi, (N, 1)).astype(np.float32), integrator=integ, wkind=L.W_CSR,
                 cfn=L.CF_KURAMOTO, cparams=(G,), indptr=ip, indices=ix, weights=w, edge_delays=d,
                 node_param_idx=0, node_param=omega)
    target = rng.uniform(0.2, 0.8, T).astype(np.float32)
    return I, L.LOSS_ORDER, target


def g2do_fc_case(N=4, T=40000, seed=11, dt=1.0, period=10, G=0.1):
    """40 s of stochastic G2DO: long enough that FC of the BOLD series is not
    pinned at 1 by the haemodynamic onset (entries ~0.97-0.99, unclamped)."""
    rng = np.random.default_rng(seed)
    ip, ix, w, d = _net(rng, N, 0.6, 1, 8)
    x0 = rng.normal(0, 1.0, (N, 2)).astype(np.float32)
    a = rng.uniform(-1.5, 1.5, N).astype(np.float32)
    I = Instance(model=L.G2DO, n_nodes=N, n_steps=T, dt=dt, params=DEFAULTS[L.G2DO].copy(), x0=x0,
                 integrator=L.HEUN_STOCH, nsig=np.full(2, 1e-3, np.float32),
                 noise=rng.standard_normal((T, N, 2)).astype(np.float32), wkind=L.W_CSR, cfn=L.CF_LINEAR,
                 cparams=(G, 0.0), indptr=ip, indices=ix, weights=w, edge_delays=d, node_param_idx=2,
                 node_param=a, bold_period=period)
    A = rng.uniform(0.5, 0.95, (N, N)).astype(np.float32)
    target = ((A + A.T) / 2).astype(np.float32)
    np.fill_diagonal(target, 1.0)
    return I, L.LOSS_FC, target


# Per model: time step (ms), steps, BOLD period -- horizons long enough for the
# BOLD series to depend measurably on the network (the haemodynamic response
# builds over hundreds of ms).
BOLD_SETUP = {L.G2DO: (1.0, 2000, 5), L.MPR: (0.1, 2000, 10), L.JANSEN_RIT: (0.1, 2000, 10),
              L.WILSON_COWAN: (0.1, 2000, 10), L.RWW: (0.1, 2000, 10), L.KURAMOTO: (0.1, 2000, 10)}


def bold_mse_case(model, orc, N=5, seed=3, integ=L.HEUN, dmin=1, dmax=10, G=0.05, node_param_idx=-1,
                  noise=False, T=None):
    """BOLD-MSE loss against 0.8 x the run's own FP64 BOLD, so residuals are on
    the scale of the signal."""
    rng = np.random.default_rng(seed)
    ip, ix, w, d = _net(rng, N, 0.5, dmin, dmax)
    nv = L.MODEL_INFO[model][0]
    dt, T0, period = BOLD_SETUP[model]
    T = T0 if T is None else T
    if model == L.MPR:
        x0 = np.stack([np.full(N, 0.02), np.full(N, -1.6)], 1) + rng.normal(0, 0.005, (N, 2))
    elif model in (L.WILSON_COWAN, L.RWW):
        x0 = rng.uniform(0.1, 0.5, (N, nv))
    elif model == L.KURAMOTO:
        x0 = rng.uniform(0, 6, (N, 1))
    elif model == L.G2DO:
        x0 = rng.normal(0, 1.0, (N, nv))
    else:
        x0 = rng.normal(0, 0.1, (N, nv))
    cfn, cp = (L.CF_KURAMOTO, (G,)) if model == L.KURAMOTO else (L.CF_LINEAR, (G, 0.0))
    p = DEFAULTS[model].copy()
    kw = {}
    if node_param_idx >= 0:
        kw = dict(node_param_idx=node_param_idx,
                  node_param=(p[node_param_idx] * (1.0 + 0.05 * rng.standard_normal(N))).astype(np.float32))
    if noise:
        kw.update(nsig=np.full(nv, 1e-4, np.float32), noise=rng.standard_normal((T, N, nv)).astype(np.float32))
    I = Instance(model=model, n_nodes=N, n_steps=T, dt=dt, params=p, x0=x0.astype(np.float32),
                 integrator=integ, wkind=L.W_CSR, cfn=cfn, cparams=cp, indptr=ip, indices=ix, weights=w,
                 edge_delays=d, bold_period=period, **kw)
    o = orc.run(dataclasses.replace(I, record_bold=True, bold_tr_ms=0.0), dtype=np.float64, native_bold=True)
    target = (0.8 * o["bold"]).astype(np.float32)
    return I, L.LOSS_BOLD_MSE, target


# ------------------------------------------------------------- oracle loss --

def fc64(B):
    n = B.shape[0]
    m = B.mean(0)
    A = B - m
    sd = np.sqrt((A * A).sum(0) / n)
    sd = np.where(sd < 1e-12, 1.0, sd)
    return np.clip((A.T @ A) / n / (sd[:, None] * sd[None, :]), -1.0, 1.0)


def oracle_loss(orc, I, kind, target, dtype=np.float64):
    if kind == L.LOSS_ORDER:
        o = orc.run(dataclasses.replace(I, record_traj=True), dtype=dtype)
        th = o["traj"][:, :, 0].astype(np.float64)
        R = np.sqrt(np.cos(th).mean(1) ** 2 + np.sin(th).mean(1) ** 2)
        return float(np.mean((R - target.astype(np.float64)) ** 2))
    o = orc.run(dataclasses.replace(I, record_bold=True, bold_tr_ms=0.0), dtype=dtype, native_bold=True)
    B = o["bold"].astype(np.float64)
    if kind == L.LOSS_FC:
        return float(np.sum((fc64(B) - target.astype(np.float64)) ** 2))
    return float(np.mean((B - target.astype(np.float64)) ** 2))


def fd_grads(orc, I, kind, target, h=1e-4):
    """Central FP64 finite differences of the oracle loss in G and in each
    node's parameter (the parameters are f32 in the problem struct, so each
    step is divided by the f32 difference actually applied)."""
    def with_G(g):
        cp = list(I.cparams)
        cp[0] = g
        return dataclasses.replace(I, cparams=tuple(cp))
    g0 = float(np.float32(I.cparams[0]))
    gp, gm = np.float32(g0 + h), np.float32(g0 - h)
    dG = (oracle_loss(orc, with_G(float(gp)), kind, target) - oracle_loss(orc, with_G(float(gm)), kind, target)) / (
        float(gp) - float(gm))
    dP = None
    if I.node_param_idx >= 0:
        dP = np.zeros(I.n_nodes)
        for k in range(I.n_nodes):
            base = np.asarray(I.node_param, np.float32)
            hk = h * max(1.0, abs(float(base[k])))
            up, dn = base.copy(), base.copy()
            up[k] = np.float32(base[k] + hk)
            dn[k] = np.float32(base[k] - hk)
            lu = oracle_loss(orc, dataclasses.replace(I, node_param=up), kind, target)
            ld = oracle_loss(orc, dataclasses.replace(I, node_param=dn), kind, target)
            dP[k] = (lu - ld) / (float(up[k]) - float(dn[k]))
    return dG, dP


def rel_err(a, b):
    a = np.atleast_1d(np.asarray(a, np.float64))
    b = np.atleast_1d(np.asarray(b, np.float64))
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
