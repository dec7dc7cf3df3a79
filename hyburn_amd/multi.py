"""Several subnetworks joined by projections: the multi-subnetwork form of
BatchHybridEngine::run_sweep (src/engine/batch_engine/engine.rs:540-823),
backed by hb_run_multi.

  Subnet(model, n_nodes, params, x0)        SubnetworkConfig (config.rs:83-114)
  Link(src, tgt, weights, coupling_fn, coupling_params, cvar_map, delay)
                                            ProjectionConfig: weights a float
                                            (Scalar: w * mean over source
                                            nodes) or a [n_tgt][n_src] array;
                                            delay = the first of its delays
  run_multi(subnets, links, dt, n_steps, integrator, hybrid_integrator,
            sweep=(sub_idx, param_idx, values))
      -> MultiResult(tavg, final_states, kernel_ms), one [n_sweep][n_nodes][nvar]
         array per subnetwork, as BatchSweepResult holds them.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib as L


@dataclass
class Subnet:
    model: int
    n_nodes: int
    params: Sequence[float]
    x0: Optional[np.ndarray] = None          # [n_nodes][nvar]; zeros by default


@dataclass
class Link:
    src: int
    tgt: int
    weights: object = 0.0                    # float (Scalar) or [n_tgt][n_src] array
    coupling_fn: int = L.CF_LINEAR
    coupling_params: Sequence[float] = (1.0,)
    cvar_map: Sequence[tuple] = ((0, 0),)
    delay: int = 0
    edge_delays: Optional[Sequence[int]] = None   # CSR weights (indptr, indices, data): per-edge delays (plain)


@dataclass
class MultiResult:
    tavg: list = field(default_factory=list)
    final_states: list = field(default_factory=list)
    kernel_ms: float = 0.0
    hist: Optional[np.ndarray] = None     # [n_sweep][total nodes][H][2] by step % H (delayed runs)
    trajectories: Optional[list] = None   # per subnetwork [n_steps][n_sweep][n_nodes][nvar]
    bold: Optional[np.ndarray] = None     # subnetwork 0: [n_sweep][rows][n_nodes] ([rows][n_nodes] native: bold_mean)
    bold_state: Optional[np.ndarray] = None   # bold_mean: the monitor's s, f, v, q per node after the run
    step_end: int = 0                     # the engine's step counter after the run


STIM_PATTERNS = {"impulse": 0, "step": 1, "sinusoid": 2, "pulse_train": 3}


def horizon(links: Sequence[Link], max_delay: int = 0) -> int:
    """History length: max delay + 1, 0 without delays (engine.rs:253-257);
    per-edge CSR delays count, and such a projection always has a history."""
    d = max([max_delay] + [L_.delay for L_ in links] +
            [int(np.max(L_.edge_delays)) for L_ in links if L_.edge_delays is not None and len(L_.edge_delays)])
    per_edge = any(L_.edge_delays is not None for L_ in links)
    return d + 1 if d > 0 else (1 if per_edge else 0)


def parse_cvar_map(s: str) -> list:
    """'0:0,1:1' -> [(0, 0), (1, 1)] (the ProjectionConfig.cvar_map string)."""
    return [tuple(int(v) for v in part.split(":")) for part in s.split(",") if part.strip()]


def cartesian_points(dims: Sequence[tuple]) -> list:
    """SweepSpec::cartesian_product (engine.rs:69-90): one [(sub, param, value), ...]
    per point, the last dimension varying fastest."""
    import itertools
    return [[(d[0], d[1], float(v)) for d, v in zip(dims, combo)]
            for combo in itertools.product(*[list(d[2]) for d in dims])]


def cartesian_tables(subnets: Sequence[Subnet], dims: Sequence[tuple]):
    """hb_multi_spec point_mode / point_values / point_params of a Cartesian
    sweep: per subnetwork, 1 for a single dimension on its native index, 0
    for a single other index (the batch dfun ignores it) or none, 2 for two
    or more dimensions (per-point parameter vectors, out-of-range indices
    dropped as set_model_params drops them)."""
    pts = cartesian_points(dims)
    S = len(subnets)
    per_sub = [[d for d in dims if d[0] == q] for q in range(S)]
    modes = []
    for q, ds in enumerate(per_sub):
        if len(ds) == 1 and ds[0][1] == L.SWEEP_INDEX[subnets[q].model]:
            modes.append(1)
        elif len(ds) >= 2:
            modes.append(2)
        else:
            modes.append(0)
    pvals = np.zeros((len(pts), S), np.float32)
    pparams = np.zeros((len(pts), S, L.MAX_PARAMS), np.float32)
    for b, pt in enumerate(pts):
        for q in range(S):
            base = np.asarray(subnets[q].params, np.float32)
            pparams[b, q, :len(base)] = base
        for (sub, pidx, v) in pt:
            if not 0 <= sub < S:
                raise L.HyburnError(L.HB_INVALID_CONFIG, "sweep subnetwork index out of range")
            if modes[sub] == 1:
                pvals[b, sub] = v
            elif modes[sub] == 2 and pidx < len(subnets[sub].params):
                pparams[b, sub, pidx] = v
    return modes, pvals, pparams


def run_multi(subnets: Sequence[Subnet], links: Sequence[Link], *, dt: float, n_steps: int,
              integrator: int = L.HEUN, hybrid_integrator: bool = False, n_sweep: int = 1,
              sweep: Optional[tuple] = None, max_delay: int = 0, nsig=None, noise=None,
              noise_seed: int = 0, step0: int = 0, state: Optional[list] = None,
              hist: Optional[np.ndarray] = None, plain: bool = False, stimuli: Sequence[tuple] = (),
              dt_ms: Optional[float] = None, record_trajectory: bool = False, bold_period: int = 0,
              bold_tr_ms: float = 0.0, device: int = 0, sweep_spec: Optional[Sequence[tuple]] = None,
              bold_mean: bool = False, bold_state: Optional[np.ndarray] = None) -> MultiResult:
    """max_delay: the largest entry of any projection's delay list, which sets
    the history horizon (engine.rs:253-257); each link's own delay counts too.
    Stochastic integrators: nsig per variable (a scalar is repeated over the
    first subnetwork's nvar, as NsigConfig::to_vec does) and the raw N(0,1)
    draws, one [n_steps][n_nodes][nvar] array per subnetwork shared by all
    sweep points (drawn from numpy's PCG64 seeded with noise_seed when not
    given; the reference draws from an unseeded thread_rng).
    Resume: step0 is the engine's step counter at the start, state one
    [n_sweep][n_nodes][nvar] array per subnetwork, hist the history returned
    by the previous run (MultiResult.hist).
    plain: the plain HybridEngine's semantics (true-ring delays, Scalar
    weights as w * I, stimuli); stimuli: (target, pattern, params) with
    pattern one of impulse, step, sinusoid, pulse_train (stimulus.rs), at
    times step * dt_ms (default dt).
    bold_period > 0: a BOLD monitor on subnetwork 0, one series per sweep
    point, windows of bold_period steps resampled to bold_tr_ms (<= 0: the
    native window samples), bold_monitor.rs:72-227. The reference's batch
    engine instead feeds one monitor the mean of var0 over the sweep points
    (engine.rs:785-801); the Balloon-Windkessel model is nonlinear, so with
    n_sweep > 1 these per-point series are not the reference's single series
    (bold_mean=True gives that monitor).
    bold_mean: one monitor fed var0 of subnetwork 0 averaged over the sweep
    points in point order (engine.rs:785-801); bold is then this run's native
    window samples [rows][n_nodes], the partial window flushed, and
    MultiResult.bold_state its Balloon-Windkessel state [n_nodes][4] (s, f, v,
    q), which bold_state= carries into the next run (None: from rest).
    sweep_spec: run_sweep_multi's Cartesian product (engine.rs:56-91,
    862-1348), dimensions (sub_idx, param_idx, values), the last varying
    fastest; n_sweep is the product of their lengths. Per subnetwork, one
    dimension on the model's batch-native index is the swept form, one on
    any other index is ignored, and two or more set every point's parameter
    vector (dfun.rs:38-46, 99-133). It replaces sweep."""
    keep = []
    subs = (L.HbSubnetwork * len(subnets))()
    for k, S in enumerate(subnets):
        nv = L.MODEL_INFO[S.model][0]
        p = np.ascontiguousarray(S.params, np.float32)
        x0 = np.zeros((S.n_nodes, nv), np.float32) if S.x0 is None else np.ascontiguousarray(S.x0, np.float32)
        if x0.shape != (S.n_nodes, nv):
            raise ValueError(f"subnetwork {k}: x0 must be [{S.n_nodes}][{nv}]")
        keep += [p, x0]
        subs[k] = L.HbSubnetwork(S.model, S.n_nodes, len(p), p.ctypes.data_as(L.fp), x0.ctypes.data_as(L.fp))
    projs = (L.HbMultiProjection * max(1, len(links)))()
    for k, R in enumerate(links):
        P = projs[k]
        P.src, P.tgt, P.cfn, P.delay = R.src, R.tgt, R.coupling_fn, R.delay
        if np.isscalar(R.weights):
            P.wkind, P.scalar_w = L.W_SCALAR_MEAN, float(R.weights)
        elif isinstance(R.weights, tuple) and len(R.weights) == 3:   # CSR by target row
            ipt, ind, dat = (np.ascontiguousarray(R.weights[0], np.int32), np.ascontiguousarray(R.weights[1], np.int32),
                             np.ascontiguousarray(R.weights[2], np.float32))
            keep += [ipt, ind, dat]
            P.wkind, P.nnz = L.W_CSR, len(dat)
            P.indptr, P.indices, P.csr_w = ipt.ctypes.data_as(L.ip), ind.ctypes.data_as(L.ip), dat.ctypes.data_as(L.fp)
            if R.edge_delays is not None:
                ed = np.ascontiguousarray(R.edge_delays, np.int32)
                keep.append(ed)
                P.edge_delays = ed.ctypes.data_as(L.ip)
        else:
            W = np.ascontiguousarray(R.weights, np.float32)
            if W.shape != (subnets[R.tgt].n_nodes, subnets[R.src].n_nodes):
                raise ValueError(f"projection {k}: dense weights must be [n_tgt][n_src]")
            keep.append(W)
            P.wkind, P.dense_w = L.W_DENSE, W.ctypes.data_as(L.fp)
        cp = list(R.coupling_params)
        P.n_cparams = len(cp)
        for j, v in enumerate(cp[:8]):
            P.cparams[j] = v
        cm = list(R.cvar_map)
        P.n_map = len(cm)
        for j, (a, b) in enumerate(cm[:4]):
            P.map_src[j], P.map_tgt[j] = a, b
    spec = L.HbMultiSpec()
    spec.n_subnets, spec.subnets = len(subnets), subs
    spec.n_projections, spec.projections = len(links), projs
    spec.dt, spec.integrator, spec.hybrid = dt, integrator, int(bool(hybrid_integrator))
    spec.n_sweep, spec.n_steps, spec.device, spec.max_delay = n_sweep, n_steps, device, max_delay
    spec.sweep_sub, spec.sweep_param = -1, 0
    if sweep is not None:
        sub_idx, param_idx, values = sweep
        vals = np.ascontiguousarray(values, np.float32)
        if len(vals) != n_sweep:
            raise ValueError("number of sweep values must equal n_sweep")
        keep.append(vals)
        spec.sweep_sub, spec.sweep_param, spec.sweep_values = sub_idx, param_idx, vals.ctypes.data_as(L.fp)
    if sweep_spec is not None:
        modes, pvals, pparams = cartesian_tables(subnets, sweep_spec)
        if pvals.shape[0] != n_sweep:
            raise ValueError("n_sweep must equal the product of the sweep dimensions' lengths")
        keep += [pvals, pparams]
        spec.sweep_sub = -1
        for q, m in enumerate(modes):
            spec.point_mode[q] = m
        spec.point_values, spec.point_params = pvals.ctypes.data_as(L.fp), pparams.ctypes.data_as(L.fp)
    if integrator in (L.EULER_STOCH, L.HEUN_STOCH, L.RK4_STOCH):
        nv0 = L.MODEL_INFO[subnets[0].model][0]
        ns = np.atleast_1d(np.asarray(0.0 if nsig is None else nsig, np.float32))
        if ns.size == 1:
            ns = np.repeat(ns, nv0)
        if noise is None:
            rng = np.random.default_rng(noise_seed)
            noise = [rng.standard_normal((n_steps, S.n_nodes, L.MODEL_INFO[S.model][0])).astype(np.float32)
                     for S in subnets]
        z = np.concatenate([np.ascontiguousarray(a, np.float32).ravel() for a in noise])
        keep += [ns, z]
        spec.n_nsig, spec.nsig, spec.noise = len(ns), ns.ctypes.data_as(L.fp), z.ctypes.data_as(L.fp)
    sizes = [n_sweep * S.n_nodes * L.MODEL_INFO[S.model][0] for S in subnets]
    spec.step0 = int(step0)
    spec.plain, spec.dt_ms = int(bool(plain)), float(dt if dt_ms is None else dt_ms)
    if stimuli:
        st_arr = (L.HbStimulus * len(stimuli))()
        for k, (target, pattern, params) in enumerate(stimuli):
            pid = STIM_PATTERNS[pattern.lower()] if isinstance(pattern, str) else int(pattern)
            st_arr[k].target, st_arr[k].pattern, st_arr[k].n_params = target, pid, len(params)
            for j, v in enumerate(list(params)[:4]):
                st_arr[k].params[j] = v
        keep.append(st_arr)
        spec.n_stimuli, spec.stimuli = len(stimuli), st_arr
    if state is not None:
        st = np.concatenate([np.ascontiguousarray(a, np.float32).ravel() for a in state])
        if st.size != sum(sizes):
            raise ValueError("state must hold one [n_sweep][n_nodes][nvar] array per subnetwork")
        keep.append(st)
        spec.state_in = st.ctypes.data_as(L.fp)
    H = horizon(links, max_delay)
    n_tot = sum(S.n_nodes for S in subnets)
    hist_out = np.zeros((n_sweep, n_tot, H, 2), np.float32) if H else None
    if H:
        spec.hist_out = hist_out.ctypes.data_as(L.fp)
        if hist is not None:
            hi = np.ascontiguousarray(hist, np.float32)
            if hi.shape != hist_out.shape:
                raise ValueError(f"hist must be {hist_out.shape}")
            keep.append(hi)
            spec.hist_in = hi.ctypes.data_as(L.fp)
    bold = bstate = None
    if bold_period > 0 and bold_mean:
        rows = int(L.lib().hb_bold_rows(n_steps, bold_period, spec.dt_ms, 0.0))
        bold = np.zeros((rows, subnets[0].n_nodes), np.float32)
        bstate = np.zeros((subnets[0].n_nodes, 4), np.float32)
        bstate[:, 1:] = 1.0
        if bold_state is not None:
            bs = np.asarray(bold_state, np.float32)
            if bs.shape != bstate.shape:
                raise ValueError(f"bold_state must be {bstate.shape}")
            bstate[:] = bs
        spec.bold_period, spec.bold, spec.bold_mean = bold_period, bold.ctypes.data_as(L.fp), 1
        spec.bold_state = bstate.ctypes.data_as(L.fp)
    elif bold_period > 0:
        rows = int(L.lib().hb_bold_rows(n_steps, bold_period, spec.dt_ms, bold_tr_ms))
        bold = np.zeros((n_sweep, rows, subnets[0].n_nodes), np.float32)
        spec.bold_period, spec.bold_tr_ms, spec.bold = bold_period, bold_tr_ms, bold.ctypes.data_as(L.fp)
    traj = np.zeros(sum(sizes) * n_steps, np.float32) if record_trajectory else None
    if traj is not None:
        spec.trajectory = traj.ctypes.data_as(L.fp)
    tavg = np.zeros(sum(sizes), np.float32)
    fin = np.zeros(sum(sizes), np.float32)
    spec.tavg, spec.final_state = tavg.ctypes.data_as(L.fp), fin.ctypes.data_as(L.fp)
    L.check(L.lib().hb_run_multi(spec))
    out = MultiResult(kernel_ms=spec.kernel_ms, hist=hist_out, step_end=int(step0) + int(n_steps), bold=bold,
                      bold_state=bstate)
    if traj is not None:
        out.trajectories, o = [], 0
        for S, n in zip(subnets, sizes):
            out.trajectories.append(traj[o:o + n * n_steps].reshape(n_steps, n_sweep, S.n_nodes,
                                                                    L.MODEL_INFO[S.model][0]).copy())
            o += n * n_steps
    o = 0
    for S, n in zip(subnets, sizes):
        shape = (n_sweep, S.n_nodes, L.MODEL_INFO[S.model][0])
        out.tavg.append(tavg[o:o + n].reshape(shape).copy())
        out.final_states.append(fin[o:o + n].reshape(shape).copy())
        o += n
    return out
