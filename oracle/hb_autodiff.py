"""Differentiable FP64 restatement of one instance's forward graph in torch.

TEST INFRASTRUCTURE ONLY: the gradient reference for the HIP adjoint (the
counterpart of differentiating the reference's graph with Burn's autodiff
backend). torch autograd differentiates this restatement; it shares no code
with the hand-written adjoint kernels. The forward follows the FP64 twin of
the C oracle (hb_oracle_body.h), which restates the reference:

  model constants / dfuns   src/engine/batch_engine/dfun.rs:271-516
  clamps                    dfun.rs:170-239
  Euler/Heun/RK4 (+ noise)  engine.rs:676-773, integrator.rs:157-315
  per-edge delays           src/engine/runtime.rs:162-204
  coupling pre/post         src/engine/coupling.rs:204-399
  BOLD                      src/engine/bold.rs:106-166, bold_monitor.rs:72-120
  FC                        src/sbi/features/fc.rs:11-57

Scope: CSR weights with per-edge delays (or none), Linear / ScaledLinear /
Kuramoto coupling, the losses of the adjoint plan (FC of BOLD, BOLD MSE,
Kuramoto order-parameter MSE).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from hyburn_amd import _lib as HL

EPS32 = float(np.float32(1.1920929e-07))
PI = math.pi


def model_consts(model: int, p: torch.Tensor) -> list:
    """p [..., n_params] float64 -> list of constants (tensors), the oracle's
    mconst_init in double."""
    P = [p[..., i] for i in range(p.shape[-1])]
    t = lambda v: torch.as_tensor(v, dtype=p.dtype)
    if model == HL.G2DO:
        tau = torch.where(P[0] > EPS32, P[0], t(EPS32))
        return dict(i=P[1], a=P[2], b=P[3], c=P[4], e=P[6], f=P[7], g=P[8], alpha=P[9], beta=P[10],
                    gamma=P[11], dtau=P[5] * tau, dovt=P[5] / tau)
    if model == HL.MPR:
        tau = torch.where(P[0] > EPS32, P[0], t(EPS32))
        return dict(inv_tau=1.0 / tau, dr_c=P[1] / (PI * tau), pp=PI * PI * tau * tau, jtau=P[3] * tau,
                    eta_i=P[2] + P[4], cr=P[5], cv=P[6])
    if model == HL.JANSEN_RIT:
        J = P[7]
        return dict(v0=P[4], r=P[6], nu2=2.0 * P[5], Aa=P[0] * P[2], a2=2.0 * P[2], aa=P[2] * P[2],
                    a1J=P[8] * J, a2J=P[9] * J, a3J=P[10] * J, a4J=P[11] * J, Bb=P[1] * P[3],
                    b2=2.0 * P[3], bb=P[3] * P[3], mu=P[12])
    if model == HL.WILSON_COWAN:
        te = torch.where(P[4] > EPS32, P[4], t(EPS32))
        ti = torch.where(P[5] > EPS32, P[5], t(EPS32))
        d = dict(cee=P[0], cei=P[1], cie=P[2], cii=P[3], inv_te=1.0 / te, inv_ti=1.0 / ti, ae=P[6], be=P[7],
                 ce=P[8], the=P[9], ai=P[10], bi=P[11], ci=P[12], thi=P[13], re=P[14], ri=P[15], ke=P[16],
                 ki=P[17], p=P[18], q=P[19], ale=P[20], ali=P[21])
        d["off_e"] = 1.0 / (1.0 + torch.exp(P[6] * P[7]))
        d["off_i"] = 1.0 / (1.0 + torch.exp(P[10] * P[11]))
        return d
    if model == HL.KURAMOTO:
        return dict(omega=P[0])
    return dict(a=P[0], b=P[1], d=P[2], gamma=P[3], inv_taus=1.0 / P[4], wjn=P[5] * P[6], jn=P[6], io=P[7])


def dfun(model: int, k: dict, x: torch.Tensor, c: torch.Tensor) -> torch.Tensor:
    """x [N, NV], c [N, NC] -> dx [N, NV] (oracle dfun, non-swept form)."""
    if model == HL.G2DO:
        v, w = x[:, 0], x[:, 1]
        iext = c[:, 0] * k["gamma"] + k["gamma"] * k["i"]
        v2 = v * v
        v3 = v * v2
        d0 = ((((w * k["alpha"] + iext) - v3 * k["f"]) + v2 * k["e"]) + v * k["g"]) * k["dtau"]
        d1 = (((v * k["b"] + v2 * k["c"]) - w * k["beta"]) + k["a"]) * k["dovt"]
        return torch.stack([d0, d1], 1)
    if model == HL.MPR:
        r, v = x[:, 0], x[:, 1]
        d0 = ((r * v) * 2.0 + k["dr_c"]) * k["inv_tau"]
        dv = ((((v * v - r * r * k["pp"]) + k["eta_i"]) + r * k["jtau"]) + c[:, 0] * k["cr"]) + c[:, 1] * k["cv"]
        return torch.stack([d0, dv * k["inv_tau"]], 1)
    if model == HL.JANSEN_RIT:
        sig = lambda u: k["nu2"] / (torch.exp((-u + k["v0"]) * k["r"]) + 1.0)
        s12 = sig(x[:, 1] - x[:, 2])
        s01 = sig(x[:, 0] * k["a1J"])
        s03 = sig(x[:, 0] * k["a3J"])
        mut = (s01 * k["a2J"] + k["mu"]) + c[:, 0]
        d3 = (s12 * k["Aa"] - x[:, 3] * k["a2"]) - x[:, 0] * k["aa"]
        d4 = (mut * k["Aa"] - x[:, 4] * k["a2"]) - x[:, 1] * k["aa"]
        d5 = ((s03 * k["a4J"]) * k["Bb"] - x[:, 5] * k["b2"]) - x[:, 2] * k["bb"]
        return torch.stack([x[:, 3], x[:, 4], x[:, 5], d3, d4, d5], 1)
    if model == HL.WILSON_COWAN:
        e, i = x[:, 0], x[:, 1]
        xe = ((e * k["cee"] - i * k["cei"]) - k["the"]) + c[:, 0]
        xe = (xe + k["p"]) * k["ale"]
        xi = (((e * k["cie"] - i * k["cii"]) + k["q"]) - k["thi"]) * k["ali"]
        se = (1.0 / (torch.exp((xe - k["be"]) * (-k["ae"])) + 1.0) - k["off_e"]) * k["ce"]
        si = (1.0 / (torch.exp((xi - k["bi"]) * (-k["ai"])) + 1.0) - k["off_i"]) * k["ci"]
        d0 = (-e + (e * (-k["re"]) + k["ke"]) * se) * k["inv_te"]
        d1 = (-i + (i * (-k["ri"]) + k["ki"]) * si) * k["inv_ti"]
        return torch.stack([d0, d1], 1)
    if model == HL.KURAMOTO:
        return (c[:, 0] + k["omega"]).unsqueeze(1)
    s = x[:, 0]
    xx = (s * k["wjn"] + k["io"]) + c[:, 0] * k["jn"]
    axb = xx * k["a"] - k["b"]
    h = axb / (1.0 - torch.exp(axb * (-k["d"])))
    return ((-s) * k["inv_taus"] + ((1.0 - s) * h) * k["gamma"]).unsqueeze(1)


def clamp(model: int, x: torch.Tensor) -> torch.Tensor:
    z = torch.zeros_like(x[:, 0])
    if model == HL.WILSON_COWAN:
        cols = [torch.where(x[:, v] < 0, z, torch.where(x[:, v] > 1, z + 1.0, x[:, v])) for v in range(2)]
        return torch.stack(cols, 1)
    if model == HL.MPR:
        return torch.stack([torch.where(x[:, 0] < 0, z, x[:, 0]), x[:, 1]], 1)
    if model == HL.RWW:
        return torch.where(x < 0, torch.zeros_like(x), torch.where(x > 1, torch.ones_like(x), x))
    if model == HL.KURAMOTO:
        tp = 2.0 * PI
        return x - torch.floor(x / tp).detach() * tp
    return x


def bold_step(z_raw, s, f, v, q, dt_s):
    """bold.rs:106-150 (oracle bold_step, double)."""
    kappa, gamma, tau_inv = 1.0 / float(np.float32(1.54)), 1.0 / float(np.float32(0.50)), 1.0 / float(np.float32(1.02))
    ialpha, e0 = 1.0 / float(np.float32(0.32)), float(np.float32(0.34))
    lo = float(np.float32(1e-6))
    z = torch.where(z_raw > 0, z_raw, torch.zeros_like(z_raw))
    fs = torch.where(f > lo, f, torch.full_like(f, lo))
    vs = torch.where(v > lo, v, torch.full_like(v, lo))
    ds = (z - kappa * s) - gamma * (fs - 1.0)
    pv = torch.pow(vs, ialpha)
    dv = tau_inv * (fs - pv)
    dq = tau_inv * ((fs * (1.0 - torch.pow(torch.as_tensor(1.0 - e0, dtype=fs.dtype), 1.0 / fs))) / e0 - pv * (q / vs))
    sn, fn, vn, qn = s + dt_s * ds, fs + dt_s * s, vs + dt_s * dv, q + dt_s * dq
    s = torch.where(sn > 0, sn, torch.zeros_like(sn))
    f = torch.where(fn > lo, fn, torch.full_like(fn, lo))
    v = torch.where(vn > lo, vn, torch.full_like(vn, lo))
    q = torch.where(qn > 0, qn, torch.zeros_like(qn))
    vv = torch.where(v > lo, v, torch.full_like(v, lo))
    sig = 4.0 * ((7.0 * (1.0 - q) + 2.0 * (1.0 - q / vv)) + 2.0 * (1.0 - vv))
    return sig, s, f, v, q


def fc(X: torch.Tensor) -> torch.Tensor:
    """Pearson FC of [T, N] with population std and clamp [-1, 1] (fc.rs:11-57), double."""
    n = X.shape[0]
    m = X.mean(0)
    A = X - m
    sd = torch.sqrt((A * A).sum(0) / n)
    sd = torch.where(sd < 1e-12, torch.ones_like(sd), sd)
    r = (A.T @ A) / n / (sd[:, None] * sd[None, :])
    return torch.clamp(r, -1.0, 1.0)


def forward(inst, G: torch.Tensor, node_p: torch.Tensor | None = None, x0: torch.Tensor | None = None,
            loss_kind: int | None = None, target=None, record: bool = False):
    """Run one hyburn_amd.Instance in FP64 with autograd. G = cparams[0] (a
    0-d tensor), node_p = the per-node values of params[node_param_idx] (or
    None), x0 = initial state [N, NV] (or None: the instance's). Returns the
    loss (0-d tensor), plus a dict of traces when record."""
    dtp = torch.float64
    N, T, model = inst.n_nodes, inst.n_steps, inst.model
    nv, nc = HL.MODEL_INFO[model][0], HL.MODEL_INFO[model][1]
    pbase = torch.as_tensor(np.asarray(inst.params, np.float32).astype(np.float64))
    if inst.node_param_idx >= 0:
        vals = node_p if node_p is not None else torch.as_tensor(np.asarray(inst.node_param, np.float64))
        cols = [pbase[i].expand(N) if i != inst.node_param_idx else vals for i in range(len(pbase))]
        P = torch.stack(cols, 1)
    else:
        P = pbase.expand(N, len(pbase))
    k = model_consts(model, P)
    x = x0 if x0 is not None else torch.as_tensor(np.asarray(inst.x0, np.float32).astype(np.float64).reshape(N, nv))
    dt = float(np.float32(inst.dt))
    integ = inst.integrator
    stoch = integ >= HL.EULER_STOCH
    if stoch:
        noise = torch.as_tensor(np.asarray(inst.noise, np.float32).astype(np.float64).reshape(T, N, nv))
        nsig = np.asarray(inst.nsig, np.float32).astype(np.float64)
        nscale = torch.as_tensor(np.sqrt(2.0 * nsig) * np.sqrt(dt))
    has_c = inst.wkind == HL.W_CSR
    kur = inst.cfn == HL.CF_KURAMOTO
    cp = list(inst.cparams) + [0.0] * 8
    b = float(np.float32(cp[1])) if len(inst.cparams) > 1 else 0.0
    if has_c:
        ip = np.asarray(inst.indptr)
        tgt = np.repeat(np.arange(N), np.diff(ip))
        src = np.asarray(inst.indices)
        w = torch.as_tensor(np.asarray(inst.weights, np.float32).astype(np.float64))
        d = np.asarray(inst.edge_delays) if inst.edge_delays is not None else np.zeros(len(src), np.int64)
        groups = {int(dv): np.nonzero(d == dv)[0] for dv in np.unique(d)}
    hist = []
    traj = []
    bold_rows = []
    P_b = inst.bold_period if loss_kind in (HL.LOSS_FC, HL.LOSS_BOLD_MSE) else 0
    if P_b:
        bs = torch.zeros(N, dtype=dtp); bf = torch.ones(N, dtype=dtp)
        bv = torch.ones(N, dtype=dtp); bq = torch.ones(N, dtype=dtp)
        bacc = torch.zeros(N, dtype=dtp); bcnt = 0
    dt_ms = inst.dt if inst.dt_ms is None else inst.dt_ms
    for t in range(T):
        xc = x[:, :nc]
        pre = torch.cat([torch.sin(xc), torch.cos(xc)], 1) if kur else xc
        hist.append(pre)
        if has_c:
            npre = pre.shape[1]
            g = torch.zeros(N, npre, dtype=dtp)
            for dv, idx in groups.items():
                if t == 0 or dv == 0:
                    srcv = pre
                elif dv > t:
                    continue
                else:
                    srcv = hist[t - dv]
                contrib = srcv[src[idx]] * w[idx][:, None]
                g = g.index_add(0, torch.as_tensor(tgt[idx]), contrib)
            if kur:
                scale = G / N
                c = ((torch.cos(xc[:, 0]) * g[:, 0] - torch.sin(xc[:, 0]) * g[:, 1]) * scale).unsqueeze(1)
            elif inst.cfn == HL.CF_SCALED_LINEAR:
                c = (g - b) * G
            else:
                c = g * G + b
        else:
            c = torch.zeros(N, nc, dtype=dtp)
        z = noise[t] * nscale if stoch else None
        k1 = dfun(model, k, x, c)
        if integ in (HL.EULER, HL.EULER_STOCH):
            xn = x + k1 * dt
            if stoch:
                xn = xn + z
            xn = clamp(model, xn)
        elif integ in (HL.HEUN, HL.HEUN_STOCH):
            s = x + k1 * dt
            if stoch:
                s = s + z
            s = clamp(model, s)
            k2 = dfun(model, k, s, c)
            xn = x + (k1 + k2) * (dt * 0.5)
            if stoch:
                xn = xn + z
            xn = clamp(model, xn)
        else:
            k2 = dfun(model, k, clamp(model, x + k1 * (dt / 2)), c)
            k3 = dfun(model, k, clamp(model, x + k2 * (dt / 2)), c)
            k4 = dfun(model, k, clamp(model, x + k3 * dt), c)
            xn = x + (((k1 + k2 * 2.0) + k3 * 2.0) + k4) * (dt / 6.0)
            if integ == HL.RK4_STOCH:
                if inst.engine == HL.ENGINE_PLAIN:
                    xn = clamp(model, xn) + z
                else:
                    xn = clamp(model, xn + z)
            else:
                xn = clamp(model, xn)
        x = xn
        traj.append(x)
        if P_b:
            bacc = bacc + x[:, 0]
            bcnt += 1
            if bcnt >= P_b or t == T - 1:
                dts = dt_ms * bcnt / 1000.0
                sig, bs, bf, bv, bq = bold_step(bacc / bcnt, bs, bf, bv, bq, dts)
                bold_rows.append(sig)
                bacc = torch.zeros(N, dtype=dtp)
                bcnt = 0
    out = {}
    if loss_kind == HL.LOSS_ORDER:
        th = torch.stack([xx[:, 0] for xx in traj], 0)        # [T, N]
        R = torch.sqrt(torch.cos(th).mean(1) ** 2 + torch.sin(th).mean(1) ** 2)
        tg = torch.as_tensor(np.asarray(target, np.float32).astype(np.float64))
        loss = ((R - tg) ** 2).mean()
        out["R"] = R
    elif loss_kind in (HL.LOSS_FC, HL.LOSS_BOLD_MSE):
        B = torch.stack(bold_rows, 0)
        out["bold"] = B
        tg = torch.as_tensor(np.asarray(target, np.float32).astype(np.float64))
        if loss_kind == HL.LOSS_FC:
            F = fc(B)
            out["fc"] = F
            loss = ((F - tg) ** 2).sum()
        else:
            loss = ((B - tg) ** 2).mean()
    else:
        loss = torch.zeros((), dtype=dtp)
    if record:
        out["traj"] = torch.stack(traj, 0)
        return loss, out
    return loss


def gradients(inst, loss_kind: int, target, grad_param: bool = True):
    """(loss, dL/dG, dL/d node parameter [N] or None, dL/dx0 [N, NV]) by autograd."""
    G = torch.tensor(float(np.float32(inst.cparams[0])) if len(inst.cparams) else 1.0, dtype=torch.float64,
                     requires_grad=True)
    nv = HL.MODEL_INFO[inst.model][0]
    x0 = torch.tensor(np.asarray(inst.x0, np.float32).astype(np.float64).reshape(inst.n_nodes, nv), requires_grad=True)
    node_p = None
    if grad_param and inst.node_param_idx >= 0:
        node_p = torch.tensor(np.asarray(inst.node_param, np.float32).astype(np.float64), requires_grad=True)
    loss = forward(inst, G, node_p, x0, loss_kind, target)
    loss.backward()
    return (float(loss.detach()), float(G.grad), None if node_p is None else node_p.grad.numpy().copy(),
            x0.grad.numpy().copy())
