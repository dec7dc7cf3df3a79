"""ctypes wrapper of the CPU parity oracle (oracle/hb_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker or the timed CPU
baseline; the product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import itertools
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libhb_oracle.so")

fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int32)
dp = C.POINTER(C.c_double)


class OrcCoupling(C.Structure):
    _fields_ = [("cfn", C.c_int32), ("n_cparams", C.c_int32), ("cparams", C.c_float * 8),
                ("n_src", C.c_int32), ("cpu_form", C.c_int32), ("wkind", C.c_int32),
                ("scalar_w", C.c_float), ("dense_w", fp), ("indptr", ip), ("indices", ip),
                ("weights", fp), ("edge_delays", ip), ("delay_mode", C.c_int32),
                ("single_delay", C.c_int32), ("max_delay", C.c_int32)]


class OrcProblem(C.Structure):
    _fields_ = [("model", C.c_int32), ("integrator", C.c_int32), ("engine", C.c_int32),
                ("n_nodes", C.c_int32), ("n_steps", C.c_int32), ("dt", C.c_float),
                ("dt_ms", C.c_double), ("params", fp), ("sweep_form", C.c_int32),
                ("sweep_val", C.c_float), ("node_param_idx", C.c_int32), ("node_param", fp),
                ("cpl", OrcCoupling), ("nsig", fp), ("noise", fp), ("x0", fp),
                ("bold_period", C.c_int32), ("bold_cap", C.c_int32),
                ("out_tavg", C.c_void_p), ("out_final", C.c_void_p), ("out_traj", C.c_void_p),
                ("out_bold", C.c_void_p), ("out_bold_n", ip)]


_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.hb_orc_run_f32.argtypes = [C.POINTER(OrcProblem)]
        L.hb_orc_run_f64.argtypes = [C.POINTER(OrcProblem)]
        L.hb_orc_run_many_f32.argtypes = [C.POINTER(OrcProblem), C.c_int32, C.c_int32]
        L.hb_orc_dfun_f32.argtypes = [C.c_int, fp, C.c_int, C.c_float, fp, fp, fp]
        L.hb_orc_dfun_f64.argtypes = [C.c_int, fp, C.c_int, C.c_float, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.hb_orc_clamp_f32.argtypes = [C.c_int, fp]
        L.hb_orc_clamp_f64.argtypes = [C.c_int, dp]
        L.hb_orc_coupling_f32.argtypes = [C.POINTER(OrcCoupling), C.c_int, C.c_int, C.c_int, fp, fp, fp]
        L.hb_orc_coupling_f64.argtypes = [C.POINTER(OrcCoupling), C.c_int, C.c_int, C.c_int, dp, dp, dp]
        L.hb_orc_csr_matvec_f32.argtypes = [fp, ip, ip, C.c_int, fp, fp]
        L.hb_orc_pre_f32.argtypes = [C.POINTER(OrcCoupling), C.c_int, C.c_int, fp, fp]
        L.hb_orc_pre_f32.restype = C.c_int
        L.hb_orc_post_f32.argtypes = [C.POINTER(OrcCoupling), C.c_int, C.c_int, fp, fp, C.c_float, fp]
        L.hb_orc_bold_step_f32.argtypes = [C.c_int, fp, C.c_float, fp, fp, fp, fp]
        L.hb_orc_bold_step_f64.argtypes = [C.c_int] + [C.POINTER(C.c_double)] + [C.c_double] + [C.POINTER(C.c_double)] * 4
        L.hb_orc_bold_signal_f32.argtypes = [C.c_int, fp, fp, fp]
        L.hb_orc_bold_signal_f64.argtypes = [C.c_int, dp, dp, dp]
        L.hb_orc_bold_downsample.argtypes = [C.POINTER(C.c_double), fp, C.c_int, C.c_int, C.c_double,
                                             C.c_double, fp, C.c_int]
        L.hb_orc_fc.argtypes = [fp, C.c_int, C.c_int, fp]
        L.hb_orc_fcd.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, fp]
        L.hb_orc_source_time.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_int64]
        L.hb_orc_source_time.restype = C.c_int64
        L.hb_orc_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _f(a):
    return None if a is None else np.ascontiguousarray(a, np.float32)


def _p(a, t=fp):
    return None if a is None else a.ctypes.data_as(t)


def make_problem(inst, keep: list, noise=None, dtype=np.float32):
    """Translate a hyburn_amd.Instance into the oracle's problem struct. The
    coupling preprocessing (Difference rewrite) mirrors the reference's
    construction (construction.rs:694-707, engine.rs:327-350)."""
    from hyburn_amd import _lib as HL
    P = OrcProblem()
    nv = inst.nvar
    k = lambda a: (keep.append(a), a)[1]
    P.model, P.integrator, P.engine = inst.model, inst.integrator, inst.engine
    P.n_nodes, P.n_steps = inst.n_nodes, inst.n_steps
    P.dt = float(np.float32(inst.dt))
    P.dt_ms = float(inst.dt if inst.dt_ms is None else inst.dt_ms)
    P.params = _p(k(_f(inst.params)))
    P.sweep_form, P.sweep_val = inst.sweep_form, float(inst.sweep_val)
    P.node_param_idx = inst.node_param_idx
    P.node_param = _p(k(_f(inst.node_param)))
    P.nsig = _p(k(_f(inst.nsig)))
    if inst.integrator >= HL.EULER_STOCH:
        if noise is None:
            noise = inst.noise
        if noise is None:
            from hyburn_amd.synth import philox_normal
            noise = philox_normal(inst.noise_seed, inst.noise_stream, inst.n_steps * inst.n_nodes * nv)
        P.noise = _p(k(_f(noise)))
    P.x0 = _p(k(_f(np.asarray(inst.x0).reshape(inst.n_nodes, nv))))
    c = P.cpl
    cfn = inst.cfn
    cp = list(inst.cparams)[:8]
    dense = inst.dense
    wkind = inst.wkind
    scalar_w = inst.scalar_w
    if cfn == HL.CF_DIFFERENCE:
        a = cp[0] if cp else 1.0
        cfn, cp = HL.CF_LINEAR, [a, 0.0]
        if wkind in (HL.W_SCALAR_MEAN, HL.W_SCALAR_DIAG):
            scalar_w = 0.0
        elif wkind == HL.W_DENSE:
            W = np.array(dense, np.float32, copy=True)
            for i in range(W.shape[0]):
                rs = np.float32(0.0)
                for j in range(W.shape[1]):
                    rs = np.float32(rs + W[i, j])
                W[i, i] = np.float32(W[i, i] - rs)
            dense = W
    c.cfn = cfn
    c.n_cparams = len(cp)
    for i, v in enumerate(cp):
        c.cparams[i] = float(v)
    c.n_src = inst.n_nodes
    c.wkind = wkind
    c.scalar_w = float(scalar_w)
    c.dense_w = _p(k(_f(dense)))
    c.indptr = _p(k(None if inst.indptr is None else np.ascontiguousarray(inst.indptr, np.int32)), ip)
    c.indices = _p(k(None if inst.indices is None else np.ascontiguousarray(inst.indices, np.int32)), ip)
    c.weights = _p(k(_f(inst.weights)))
    per_edge = wkind == HL.W_CSR and inst.delay_mode == HL.DELAY_PER_EDGE
    if per_edge and inst.edge_delays is not None:
        c.edge_delays = _p(k(np.ascontiguousarray(inst.edge_delays, np.int32)), ip)
    dm = inst.delay_mode
    if wkind != HL.W_CSR and dm == HL.DELAY_PER_EDGE:
        dm = HL.DELAY_SINGLE
    c.cpu_form = 1 if (per_edge and inst.edge_delays is not None) else 0
    c.delay_mode = dm
    c.single_delay = inst.single_delay
    c.max_delay = inst.max_delay()
    P.bold_period = inst.bold_period if inst.record_bold else 0
    return P


def run(inst, dtype=np.float32, noise=None, native_bold: bool = False):
    """Run one instance on the oracle (FP32 twin or FP64 twin).
    Returns dict(tavg, final_state, traj, bold) with BOLD resampled to TR as
    bold_monitor.rs:123-203 does (native samples when native_bold)."""
    keep: list = []
    P = make_problem(inst, keep, noise=noise)
    N, nv, T = inst.n_nodes, inst.nvar, inst.n_steps
    tavg = np.zeros((N, nv), dtype)
    fin = np.zeros((N, nv), dtype)
    traj = np.zeros((T, N, nv), dtype) if inst.record_traj else None
    nb_cap = (T // inst.bold_period + 1) if (inst.record_bold and inst.bold_period > 0) else 0
    bold = np.zeros((max(nb_cap, 1), N), dtype) if nb_cap else None
    nbn = np.zeros(1, np.int32)
    P.out_tavg = tavg.ctypes.data
    P.out_final = fin.ctypes.data
    P.out_traj = traj.ctypes.data if traj is not None else None
    P.out_bold = bold.ctypes.data if bold is not None else None
    P.bold_cap = nb_cap
    P.out_bold_n = nbn.ctypes.data_as(ip)
    f = lib().hb_orc_run_f32 if dtype == np.float32 else lib().hb_orc_run_f64
    rc = f(C.byref(P))
    if rc != 0:
        raise RuntimeError("oracle run failed")
    out = {"tavg": tavg, "final_state": fin, "traj": traj, "bold": None}
    if bold is not None:
        nb = int(nbn[0])
        native = bold[:nb]
        out["bold_native"] = native
        out["bold"] = native if native_bold else bold_resample(native, inst)
    return out


def bold_times(inst, n: int) -> np.ndarray:
    """Native BOLD sample times (ms) as bold_monitor.rs:100,214 form them in f64."""
    dt_ms = inst.dt if inst.dt_ms is None else inst.dt_ms
    P, T = inst.bold_period, inst.n_steps
    dtb = dt_ms * P
    t = []
    for k in range(n):
        if k < T // P:
            t.append(float(k) * dtb + dtb)
        else:
            t.append(float(k) * dtb + dt_ms * (T % P))
    return np.array(t, np.float64)


def bold_resample(native: np.ndarray, inst) -> np.ndarray:
    n, N = native.shape
    if n == 0:
        return native
    dt_ms = inst.dt if inst.dt_ms is None else inst.dt_ms
    times = bold_times(inst, n)
    cap = n * 64 + 1
    out = np.zeros((cap, N), np.float32)
    nat = np.ascontiguousarray(native, np.float32)
    m = lib().hb_orc_bold_downsample(times.ctypes.data_as(C.POINTER(C.c_double)), _p(nat), n, N,
                                     dt_ms * inst.bold_period, inst.bold_tr_ms, _p(out), cap)
    return out[:m]


def run_many_f32(insts, n_threads: int = 0) -> float:
    """Run instances on host threads (OpenMP); returns wall seconds."""
    import time
    keep: list = []
    arr = (OrcProblem * len(insts))()
    outs = []
    for i, I in enumerate(insts):
        P = make_problem(I, keep)
        N, nv = I.n_nodes, I.nvar
        t = np.zeros((N, nv), np.float32)
        f = np.zeros((N, nv), np.float32)
        outs.append((t, f))
        P.out_tavg = t.ctypes.data
        P.out_final = f.ctypes.data
        if P.bold_period > 0:   # the BOLD monitor runs as it does on the device
            cap = I.n_steps // I.bold_period + 1
            b = np.zeros((cap, N), np.float32)
            nb = np.zeros(1, np.int32)
            keep += [b, nb]
            P.out_bold, P.bold_cap, P.out_bold_n = b.ctypes.data, cap, nb.ctypes.data_as(ip)
        else:
            P.bold_period = 0
        arr[i] = P
    t0 = time.perf_counter()
    rc = lib().hb_orc_run_many_f32(arr, len(insts), n_threads)
    dt = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError("oracle batch failed")
    return dt


def dfun(model, params, x, c, sweep_form=0, sweep_val=0.0, dtype=np.float32):
    x = np.ascontiguousarray(x, dtype)
    c = np.ascontiguousarray(c, dtype)
    p = np.ascontiguousarray(params, np.float32)
    dx = np.zeros_like(x)
    if dtype == np.float32:
        lib().hb_orc_dfun_f32(model, _p(p), sweep_form, sweep_val, _p(x), _p(c), _p(dx))
    else:
        d = C.POINTER(C.c_double)
        lib().hb_orc_dfun_f64(model, _p(p), sweep_form, sweep_val, x.ctypes.data_as(d),
                              c.ctypes.data_as(d), dx.ctypes.data_as(d))
    return dx


def clamp(model, x, dtype=np.float32):
    x = np.array(x, dtype, copy=True)
    if dtype == np.float32:
        lib().hb_orc_clamp_f32(model, _p(x))
    else:
        lib().hb_orc_clamp_f64(model, _p(x, dp))
    return x


def _cfg(cfn, cparams, cpu_form=False, n_src=1):
    cpl = OrcCoupling()
    cpl.cfn = cfn
    cpl.n_cparams = len(cparams)
    for i, v in enumerate(cparams):
        cpl.cparams[i] = v
    cpl.n_src = n_src
    cpl.cpu_form = int(cpu_form)
    return cpl


def pre(cfn, cparams, x, cpu_form=False):
    """pre() of each row of x [n][nc] (coupling.rs:204-321) -> [n][npre]."""
    x = np.ascontiguousarray(x, np.float32)
    n, nc = x.shape
    out = np.zeros(n * 2 * nc, np.float32)
    np_ = lib().hb_orc_pre_f32(C.byref(_cfg(cfn, cparams, cpu_form)), n, nc, _p(x), _p(out))
    return out[:n * np_].reshape(n, np_).copy()


def post(cfn, cparams, g, x_tgt=None, rowsum=0.0, n_src=1):
    """post_with_target() of each row of g [n][npre] (coupling.rs:323-399) -> [n][nc]."""
    g = np.ascontiguousarray(g, np.float32)
    n = g.shape[0]
    nc = g.shape[1] // 2 if cfn == 2 else (x_tgt.shape[1] if x_tgt is not None else g.shape[1])
    x_tgt = np.zeros((n, nc), np.float32) if x_tgt is None else np.ascontiguousarray(x_tgt, np.float32)
    out = np.zeros((n, nc), np.float32)
    lib().hb_orc_post_f32(C.byref(_cfg(cfn, cparams, n_src=n_src)), n, nc, _p(g), _p(x_tgt), rowsum, _p(out))
    return out


def coupling(cfn, cparams, x_src, x_tgt=None, dense=None, csr=None, n_src=None, dtype=np.float32):
    """One delay-free coupling evaluation post(W @ pre(x_src), x_tgt) (FP32, or
    the FP64 twin with dtype float64; weights stay f32 values)."""
    from hyburn_amd import _lib as HL
    x_src = np.ascontiguousarray(x_src, dtype)
    ns, nc = x_src.shape
    nt = dense.shape[0] if dense is not None else len(csr[0]) - 1
    x_tgt = np.zeros((nt, nc), dtype) if x_tgt is None else np.ascontiguousarray(x_tgt, dtype)
    cpl = OrcCoupling()
    cpl.cfn = cfn
    cpl.n_cparams = len(cparams)
    for i, v in enumerate(cparams):
        cpl.cparams[i] = v
    cpl.n_src = ns if n_src is None else n_src
    keep = []
    if dense is not None:
        d = np.ascontiguousarray(dense, np.float32)
        keep.append(d)
        cpl.wkind, cpl.dense_w = HL.W_DENSE, _p(d)
    else:
        ipt, ind, dat = (np.ascontiguousarray(csr[0], np.int32), np.ascontiguousarray(csr[1], np.int32),
                         np.ascontiguousarray(csr[2], np.float32))
        keep += [ipt, ind, dat]
        cpl.wkind, cpl.indptr, cpl.indices, cpl.weights = HL.W_CSR, _p(ipt, ip), _p(ind, ip), _p(dat)
    out = np.zeros((nt, nc), dtype)
    if dtype == np.float32:
        lib().hb_orc_coupling_f32(C.byref(cpl), nt, ns, nc, _p(x_src), _p(x_tgt), _p(out))
    else:
        lib().hb_orc_coupling_f64(C.byref(cpl), nt, ns, nc, _p(x_src, dp), _p(x_tgt, dp), _p(out, dp))
    return out


def fc(traj):
    t = np.ascontiguousarray(traj, np.float32)
    T, N = t.shape
    out = np.zeros((N, N), np.float32)
    lib().hb_orc_fc(_p(t), T, N, _p(out))
    return out


def fcd(traj, window_len, stride):
    """FCD matrix of a [T][N] series (fc.rs:94-179); empty (0, 0) when W <= 1."""
    t = np.ascontiguousarray(traj, np.float32)
    T, N = t.shape
    W = (T - window_len) // stride + 1 if 2 <= window_len <= T and stride >= 1 else 0
    out = np.zeros((max(W, 0), max(W, 0)), np.float32)
    rc = lib().hb_orc_fcd(_p(t), T, N, window_len, stride, _p(out))
    if rc < 0:
        raise ValueError("window_len must be in [2, n_steps] and stride > 0")
    return out if rc > 1 else np.zeros((0, 0), np.float32)


_CVAR = {0: (0,), 1: (0, 1), 2: (1,), 3: (0,), 4: (0,), 5: (0,), 6: (0,), 7: (0, 1), 8: (0, 1), 9: (0, 1),
         10: (0,), 11: (0,), 12: (0, 3, 6), 13: (0,), 14: (0,), 15: (0,), 16: (0, 1, 2, 3), 17: (0, 1, 2, 3),
         18: (0, 1, 2, 3), 19: (0,), 20: (0,), 21: (0, 3), 22: (0,), 23: (0,), 24: (3,), 25: (0, 2), 26: (0, 3),
         27: (0, 1, 4, 5)}   # CVAR of src/model/*.rs
_IDENTITY_PRE = (0, 1, 3, 4)   # Linear, Difference, Sigmoidal, ScaledLinear: pre is the identity


def _f32sum(v):
    """Sequential f32 sum (the order the restatement fixes)."""
    return np.cumsum(np.asarray(v, np.float32), dtype=np.float32)[-1] if len(v) else np.float32(0.0)


def stimulus_value(pattern, params, step, dt_ms):
    """StimulusApplier::apply (stimulus.rs:63-128): f64 time step * dt_ms."""
    t = float(step) * float(dt_ms)
    p = [float(np.float32(v)) for v in params]
    onset = p[0]
    if pattern == "impulse":
        return np.float32(p[1]) if (abs(t - onset) < 1e-6 or (onset < t <= onset + 1e-6)) else np.float32(0.0)
    if pattern == "step":
        return np.float32(p[2]) if onset <= t < p[1] else np.float32(0.0)
    if pattern == "sinusoid":
        if t < onset:
            return np.float32(0.0)
        ph = p[3] if len(p) > 3 else 0.0
        return np.float32(np.float32(p[1]) * np.float32(math.sin(2.0 * math.pi * p[2] * (t - onset) + ph)))
    if pattern == "pulse_train":
        period, width = p[2], p[3]
        if t < onset or period <= 0.0 or width <= 0.0:
            return np.float32(0.0)
        return np.float32(p[1]) if math.fmod(t - onset, period) < width else np.float32(0.0)
    raise ValueError(pattern)


def bold_series(v0, P, dt_ms, tr_ms, dtype=np.float32):
    """BOLD of a [T][N] series of variable 0 (bold_monitor.rs:72-227): windows
    of P steps averaged in f32, one Balloon-Windkessel step per window
    (dt_s = window ms / 1000), a partial last window, then TR resampling
    unless TR is <= 0 or the window length. dtype float64: the FP64 twin of
    the windows and the Balloon-Windkessel steps (the TR resampling of its
    native rows stays the f32 interpolation)."""
    from types import SimpleNamespace
    fl = dtype
    v0 = np.asarray(v0, fl)
    T, N = v0.shape
    S, F, V, Q = (np.zeros(N, fl), np.ones(N, fl), np.ones(N, fl), np.ones(N, fl))
    acc, cnt, native = np.zeros(N, fl), 0, []

    def close(cnt):
        z = (acc / fl(cnt)).astype(fl)
        sig = np.zeros(N, fl)
        if fl is np.float32:
            lib().hb_orc_bold_step_f32(N, _p(z), np.float32((dt_ms * cnt) / 1000.0), _p(S), _p(F), _p(V), _p(Q))
            lib().hb_orc_bold_signal_f32(N, _p(V), _p(Q), _p(sig))
        else:
            lib().hb_orc_bold_step_f64(N, _p(z, dp), (dt_ms * cnt) / 1000.0, _p(S, dp), _p(F, dp),
                                       _p(V, dp), _p(Q, dp))
            lib().hb_orc_bold_signal_f64(N, _p(V, dp), _p(Q, dp), _p(sig, dp))
        native.append(sig)

    for t in range(T):
        acc = (acc + v0[t]).astype(fl)
        cnt += 1
        if cnt >= P:
            close(cnt)
            acc, cnt = np.zeros(N, fl), 0
    if cnt > 0:
        close(cnt)
    nat = np.array(native, fl).reshape(-1, N)
    if tr_ms <= 0.0 or abs(tr_ms - dt_ms * P) < 1e-6:
        return nat
    return bold_resample(nat, SimpleNamespace(dt=dt_ms, dt_ms=dt_ms, bold_period=P, n_steps=T,
                                              bold_tr_ms=tr_ms)).astype(fl)


def run_multi(subnets, links, dt, n_steps, integrator, hybrid, sweep=None, n_sweep=1, max_delay=0, nsig=None,
              noise=None, plain=False, stimuli=(), dt_ms=None, bold_period=0, bold_tr_ms=0.0, dtype=np.float32,
              sweep_spec=None):
    """The batch engine's multi-subnetwork loop (engine.rs:540-823) in f32, one
    sweep point at a time: coupling of every projection from the current
    states (engine.rs:557-673), then each subnetwork's integration with that
    frozen coupling (676-773), clamp, tavg. subnets: (model, n, params, x0
    [n][nvar]); links: (src, tgt, weights float|[nt][ns], cfn, cparams,
    cvar_map, delay=0). Delayed reads use the batch engine's history window
    literally (engine.rs:253-270, 541-589): H = max delay + 1 states,
    initially [x0, 0, ..., 0], shifted left and appended each step, read at
    slot (step - d + H) % H. Stochastic forms (engine.rs:726-763): Z = raw *
    (sqrt(2 nsig_v) * sqrt(dt)) from noise[k][step] (one draw shared by every
    sweep point); Euler x + d dt + Z; Heun with Z in predictor and corrector;
    RK4 clamp(det + Z). plain: the plain HybridEngine (runtime.rs:46-501):
    delays read the state of step - d, Scalar weights are w * I
    (construction.rs:639-654), stimuli are added to coupling variable 0 of
    their target (runtime.rs:319-357), RK4-stochastic is clamp(det) + Z
    (integrator.rs:296-315). bold_period > 0: BOLD of subnetwork 0's variable
    0 per sweep point (bold_series). Returns per subnetwork tavg and final
    [n_sweep][n][nvar], plus the BOLD [n_sweep][rows][n0] when monitored.
    sweep_spec: run_sweep_multi's dimensions (sub_idx, param_idx, values),
    expanded as SweepSpec::cartesian_product (engine.rs:69-90, last fastest);
    per subnetwork the dispatch of dfun.rs:38-46: one dimension is the
    swept form on the model's native index and ignored elsewhere, two or
    more override that point's parameters (dfun_batch_multi_fallback,
    dfun.rs:99-133)."""
    from hyburn_amd import _lib as HL
    f32 = np.float32
    fl = dtype   # float64: the FP64 twin (same op order; weights, params and raw noise stay f32 values)

    def _sum(v):
        return _f32sum(v) if fl is np.float32 else np.float64(np.sum(np.asarray(v, np.float64)))
    S = len(subnets)
    info = [HL.MODEL_INFO[m] for (m, _, _, _) in subnets]
    # Difference preprocessing, engine.rs:325-366
    prep = []
    H = max([max_delay] + [l[6] if len(l) > 6 else 0 for l in links] +
            [int(max(l[2][3])) for l in links if isinstance(l[2], tuple) and len(l[2]) == 4 and len(l[2][3])])
    per_edge = any(isinstance(l[2], tuple) and len(l[2]) == 4 for l in links)
    H = H + 1 if H > 0 else (1 if per_edge else 0)
    for (src, tgt, w, cfn, cp, cmap, *rest) in links:
        delay = rest[0] if rest else 0
        if isinstance(w, tuple) and len(w) == 4:   # per-edge CSR (plain): indptr, indices, data, delays
            if cfn not in _IDENTITY_PRE or cfn == HL.CF_DIFFERENCE:
                raise NotImplementedError("per-edge links restated for identity-pre coupling functions")
            prep.append((src, tgt, w, None, cfn, list(cp), list(cmap), 0))
            continue
        cp = list(cp)
        ns, nt = subnets[src][1], subnets[tgt][1]
        W = None if np.isscalar(w) else np.array(w, np.float32)
        if plain and W is None:
            W = (np.eye(nt, ns) * np.float32(w)).astype(np.float32)
        if cfn == HL.CF_DIFFERENCE:
            a = cp[0] if cp else 1.0
            if W is None:
                w, cfn, cp = 0.0, HL.CF_LINEAR, [a, 0.0]
            elif nt == ns:
                for i in range(nt):
                    W[i, i] = f32(W[i, i] - _f32sum(W[i]))
                cfn, cp = HL.CF_LINEAR, [a, 0.0]
        prep.append((src, tgt, w, W, cfn, cp, list(cmap), delay))
    tav = [np.zeros((n_sweep, n, info[k][0]), fl) for k, (_, n, _, _) in enumerate(subnets)]
    fin = [np.zeros_like(t) for t in tav]
    dtf = fl(np.float32(dt))   # the engine's f32 dt (engine.rs:456), widened in the FP64 twin
    dth, dt6 = fl(dtf * fl(0.5)), fl(dtf / fl(6.0))
    bold_out = []
    for b in range(n_sweep):
        v0 = []
        x = [np.array(x0, fl).reshape(n, info[k][0]) for k, (_, n, _, x0) in enumerate(subnets)]
        acc = [np.zeros_like(v) for v in x]
        sw = [0] * S
        svs = [0.0] * S
        prm = [np.asarray(sp[2], np.float32) for sp in subnets]
        if sweep_spec is not None:
            combo = list(itertools.product(*[list(d[2]) for d in sweep_spec]))[b]
            for k in range(S):
                mine = [(d[1], float(v)) for d, v in zip(sweep_spec, combo) if d[0] == k]
                if len(mine) == 1 and mine[0][0] == HL.SWEEP_INDEX[subnets[k][0]]:
                    sw[k], svs[k] = 1, mine[0][1]
                elif len(mine) >= 2:
                    prm[k] = prm[k].copy()
                    for (pi, v) in mine:
                        if pi < len(prm[k]):
                            prm[k][pi] = v
        elif sweep is not None:
            sub_idx, param_idx, values = sweep
            svs = [float(values[b])] * S
            if param_idx == HL.SWEEP_INDEX[subnets[sub_idx][0]]:
                sw[sub_idx] = 1
        hist = [[x[k].copy()] + [np.zeros_like(x[k]) for _ in range(H - 1)] for k in range(S)] if H else None
        past = [dict() for _ in range(S)]
        for step in range(n_steps):
            if H:
                for k in range(S):
                    hist[k] = hist[k][1:] + [x[k].copy()]
                    past[k][step] = x[k].copy()
            cpl = [None] * S
            for (src, tgt, w, W, cfn, cp, cmap, delay) in prep:
                ms, mt = subnets[src][0], subnets[tgt][0]
                ncs, nct = info[src][1], info[tgt][1]
                ns, nt = subnets[src][1], subnets[tgt][1]
                if isinstance(w, tuple):   # runtime.rs:160-203, identity pre: sum of w_e * x_j(step - d_e)
                    ipt, ind, dat, edl = w
                    xt = x[tgt][:, :ncs] if nct == ncs else np.zeros((nt, ncs), fl)
                    out = np.zeros((nt, ncs), fl)
                    for r_ in range(nt):
                        g = np.zeros(ncs, fl)
                        for e_ in range(ipt[r_], ipt[r_ + 1]):
                            d_ = int(edl[e_])
                            if d_ == 0 or step == 0:
                                xv = x[src][ind[e_], :ncs]
                            elif d_ <= step:
                                xv = past[src][step - d_][ind[e_], :ncs]
                            else:
                                xv = np.zeros(ncs, fl)
                            g = (g + fl(dat[e_]) * xv).astype(fl)
                        out[r_] = coupling(cfn, cp, g[None, :], xt[r_:r_ + 1], dense=np.ones((1, 1), fl),
                                           n_src=ns, dtype=fl)[0]
                elif delay == 0 or step == 0:
                    xs = x[src][:, :ncs]
                elif delay <= step and plain:
                    xs = past[src][step - delay][:, :ncs]
                elif delay <= step:
                    xs = hist[src][(step - delay + H) % H][:, :ncs]
                else:
                    xs = np.zeros((ns, ncs), fl)
                xt = x[tgt][:, :ncs] if nct == ncs else np.zeros((nt, ncs), fl)
                if isinstance(w, tuple):
                    pass
                elif W is None:
                    if cfn not in _IDENTITY_PRE:
                        raise NotImplementedError("Scalar weights with a non-identity pre")
                    g = np.array([[fl(fl(_sum(xs[:, c]) / fl(ns)) * fl(np.float32(w))) for c in range(ncs)]], fl)
                    out = coupling(cfn, cp, g, xt, dense=np.ones((nt, 1), fl), n_src=ns, dtype=fl)
                else:
                    out = coupling(cfn, cp, xs, xt, dense=W, n_src=ns, dtype=fl)
                if nct == ncs and cmap == [(0, 0)]:
                    rem = out
                else:
                    rem = np.zeros((nt, nct), fl)
                    for (sc, tc) in cmap:
                        ci = _CVAR[mt].index(tc) if tc in _CVAR[mt] else 0
                        if sc < ncs and ci < nct:
                            rem[:, ci] = rem[:, ci] + out[:, sc]
                cpl[tgt] = rem if cpl[tgt] is None else (cpl[tgt] + rem).astype(fl)
            for (target, pattern, sp) in stimuli:
                if 0 <= target < S:
                    v = stimulus_value(pattern, sp, step, dt if dt_ms is None else dt_ms)
                    if v != 0.0:
                        base = cpl[target] if cpl[target] is not None else np.zeros((subnets[target][1], info[target][1]),
                                                                                     fl)
                        base = base.copy()
                        base[:, 0] = (base[:, 0] + v).astype(fl)
                        cpl[target] = base
            for k, (m, n, _, _) in enumerate(subnets):
                params, sv = prm[k], svs[k]
                c = cpl[k] if cpl[k] is not None else np.zeros((n, info[k][1]), fl)
                # hybrid_integrator: deterministic Heun for model_prefers_heun (dfun.rs:1871-1873)
                heun = (m in HL.PREFERS_HEUN) if hybrid else (integrator == HL.HEUN)
                stoch = (not (hybrid and m in HL.PREFERS_HEUN)) and integrator in (HL.EULER_STOCH, HL.HEUN_STOCH,
                                                                                   HL.RK4_STOCH)
                if stoch:
                    scale = np.array([fl(np.sqrt(fl(2.0) * fl(nsig[v]))) * fl(np.sqrt(dtf))
                                      for v in range(info[k][0])], fl)
                    Zk = (np.asarray(noise[k][step], fl) * scale).astype(fl)
                for i in range(n):
                    xi, ci = x[k][i].copy(), c[i]
                    d = dfun(m, params, xi, ci, sw[k], sv, dtype=fl)
                    if stoch:
                        z = Zk[i]
                        if integrator == HL.EULER_STOCH:
                            xn = ((xi + d * dtf) + z).astype(fl)
                        elif integrator == HL.HEUN_STOCH:
                            pr = clamp(m, ((xi + d * dtf) + z).astype(fl), fl)
                            d2 = dfun(m, params, pr, ci, sw[k], sv, dtype=fl)
                            xn = ((xi + (d + d2) * dth) + z).astype(fl)
                        else:
                            k2 = dfun(m, params, clamp(m, xi + d * dth, fl), ci, sw[k], sv, dtype=fl)
                            k3 = dfun(m, params, clamp(m, xi + k2 * dth, fl), ci, sw[k], sv, dtype=fl)
                            k4 = dfun(m, params, clamp(m, xi + k3 * dtf, fl), ci, sw[k], sv, dtype=fl)
                            det = (xi + (((d + k2 * fl(2.0)) + k3 * fl(2.0)) + k4) * dt6).astype(fl)
                            if plain:   # integrator.rs:296-315: clamp(det) + Z, no clamp after
                                x[k][i] = (clamp(m, det, fl) + z).astype(fl)
                                continue
                            xn = (det + z).astype(fl)
                        x[k][i] = clamp(m, xn, fl)
                        continue
                    if heun:
                        pr = clamp(m, (xi + d * dtf).astype(fl), fl)
                        d2 = dfun(m, params, pr, ci, sw[k], sv, dtype=fl)
                        xn = (xi + (d + d2) * dth).astype(fl)
                    elif integrator == HL.RK4:
                        k2 = dfun(m, params, clamp(m, xi + d * dth, fl), ci, sw[k], sv, dtype=fl)
                        k3 = dfun(m, params, clamp(m, xi + k2 * dth, fl), ci, sw[k], sv, dtype=fl)
                        k4 = dfun(m, params, clamp(m, xi + k3 * dtf, fl), ci, sw[k], sv, dtype=fl)
                        xn = (xi + (((d + k2 * fl(2.0)) + k3 * fl(2.0)) + k4) * dt6).astype(fl)
                    else:   # Euler (and a configured Heun under hybrid_integrator)
                        xn = (xi + d * dtf).astype(fl)
                    x[k][i] = clamp(m, xn, fl)
                acc[k] = (acc[k] + x[k]).astype(fl)
            v0.append(x[0][:, 0].copy())
        if bold_period > 0:
            bold_out.append(bold_series(np.array(v0, fl).reshape(n_steps, subnets[0][1]), bold_period,
                                        dt if dt_ms is None else dt_ms, bold_tr_ms, dtype=fl))
        inv = fl(1.0) / fl(n_steps)
        for k in range(S):
            tav[k][b] = acc[k] * inv
            fin[k][b] = x[k]
    if bold_period > 0:
        return tav, fin, np.array(bold_out, fl)
    return tav, fin


def next_power_of_2(n):
    """catch22/fft.rs:6-15."""
    p = 1
    while p < n:
        p <<= 1
    return p


EEG_BANDS = ((0.5, 4.0), (4.0, 8.0), (8.0, 13.0), (13.0, 30.0), (30.0, 80.0))   # spectral.rs:9-15


def welch_psd(y, nfft, fs, window_width):
    """welch_psd (catch22/fft.rs:120-170): segments of window_width samples
    every window_width // 2, minus the whole series' mean, zero-padded to
    next_power_of_2(max(nfft, window_width)), |DFT|^2 summed over segments for
    bins 0..nfft/2, divided by n_segments * window_width * fs. The DFT is
    numpy's FFT (the reference's radix-2 FFT computes the same transform)."""
    y = np.asarray(y, np.float64)
    n = len(y)
    m = y.sum() / n
    nf = next_power_of_2(max(nfft, window_width))
    half = nf // 2 + 1
    hop = window_width // 2
    nseg = (n - window_width) // hop + 1 if n > window_width else 1
    psd = np.zeros(half)
    for seg in range(nseg):
        start = seg * hop
        if start + window_width > n:
            break
        d = np.zeros(nf)
        d[:window_width] = y[start:start + window_width] - m
        X = np.fft.fft(d)[:half]
        psd += X.real * X.real + X.imag * X.imag
    psd /= nseg * window_width * fs
    return psd, np.arange(half) * (fs / nf)


def spectral_features(series, fs):
    """spectral_features (spectral.rs:20-99): 5 band-power fractions, then
    centroid, spread, skewness, excess kurtosis of the Welch PSD; NaN x 9 for
    fewer than 16 samples or a non-positive / non-finite total power."""
    series = np.asarray(series, np.float32)
    nan = np.full(9, np.nan, np.float32)
    if len(series) < 16:
        return nan
    nfft = next_power_of_2(len(series))
    nperseg = min(max(len(series) // 2, 16), 256)
    psd, freqs = welch_psd(series.astype(np.float64), nfft, fs, nperseg)
    total = psd.sum()
    if not (total > 0.0) or not np.isfinite(total):
        return nan
    out = [psd[(freqs >= lo) & (freqs < hi)].sum() / total for lo, hi in EEG_BANDS]
    c = (psd * freqs).sum() / total
    spread = np.sqrt((psd * (freqs - c) ** 2).sum() / total)
    skew = (psd * (freqs - c) ** 3).sum() / total / spread ** 3 if spread > 1e-12 else 0.0
    kurt = (psd * (freqs - c) ** 4).sum() / total / spread ** 4 - 3.0 if spread > 1e-12 else 0.0
    return np.array(out + [c, spread, skew, kurt], np.float32)


def source_time(mode, s, d, H):
    return int(lib().hb_orc_source_time(mode, s, d, H))
