"""Host-side mirror of hyburn's engine API for the batched integration path.

Same names, argument meaning and error classes as the reference, backed by
the HIP library through the C ABI:

  SimConfig / SweepConfig       src/config.rs:83-114, 457-503
  SweepParam / SweepPoint / SweepSpec / BatchSweepResult
                                src/engine/batch_engine/engine.rs:15-91
  BatchHybridEngine.from_config / run_sweep / run_sweep_with_trajectory /
  run_sweep_multi               engine.rs:145-149, 477-494, 845-860
  write_npy_f32                 src/io.rs:88-125
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib as L

try:  # Python 3.11+: tomllib; 3.10 image: tomli
    import tomllib as _toml
except ImportError:  # pragma: no cover
    import tomli as _toml


class SimulationError(L.HyburnError):
    pass


def _err(code: int, msg: str):
    raise L.HyburnError(code, msg)


# ---------------------------------------------------------------- config ----

@dataclass
class SubnetworkConfig:
    model: str
    nnodes: int
    params: list
    initial_state: object            # list of floats (var-major as in TOML) or a .npy path
    nmodes: int = 1


@dataclass
class ProjectionConfig:
    src: int
    tgt: int
    weights: object                   # float | [[...]] | {data, indices, indptr}
    delays: list = field(default_factory=list)
    tract_lengths: list = field(default_factory=list)
    coupling_fn: str = "Linear"
    coupling_params: list = field(default_factory=list)
    cvar_map: str = "0:0"
    conn_type: str = "all_to_all"


@dataclass
class MonitorConfig:
    monitor_type: str
    period: Optional[float] = None
    tr: Optional[float] = 2.0
    bold_period: Optional[int] = 10


@dataclass
class SimConfig:
    sim_length: float
    dt: float
    subnetworks: list
    projections: list = field(default_factory=list)
    integrator: str = "heun"
    monitors: list = field(default_factory=list)
    nsig: object = 0.0
    speed: float = 3.0
    backend: str = "ndarray"
    noise_mode: str = "reproducible"
    base_dir: str = "."
    stimuli: list = field(default_factory=list)   # [{target, temporal, params}] (config.rs:253-258)

    @classmethod
    def from_dict(cls, d: dict, base_dir: str = ".") -> "SimConfig":
        try:
            net = d["network"]
            subs = [SubnetworkConfig(model=s["model"], nnodes=int(s["nnodes"]), params=list(s["params"]),
                                     initial_state=s["initial_state"], nmodes=int(s.get("nmodes", 1)))
                    for s in net["subnetworks"]]
            projs = [ProjectionConfig(src=int(p["src"]), tgt=int(p["tgt"]), weights=p["weights"],
                                      delays=list(p.get("delays", [])), tract_lengths=list(p.get("tract_lengths", [])),
                                      coupling_fn=p.get("coupling_fn", "Linear"),
                                      coupling_params=list(p.get("coupling_params", [])),
                                      cvar_map=p.get("cvar_map", "0:0"), conn_type=p.get("conn_type", "all_to_all"))
                     for p in net.get("projections", [])]
            mons = [MonitorConfig(monitor_type=m["monitor_type"], period=m.get("period"), tr=m.get("tr", 2.0),
                                  bold_period=m.get("bold_period", 10)) for m in d.get("monitors", [])]
            return cls(sim_length=float(d["sim_length"]), dt=float(d["dt"]), subnetworks=subs, projections=projs,
                       integrator=str(d.get("integrator", "heun")).lower(), monitors=mons, nsig=d.get("nsig", 0.0),
                       speed=float(d.get("speed", 3.0)), backend=d.get("backend", "ndarray"),
                       noise_mode=d.get("noise_mode", "reproducible"), base_dir=base_dir,
                       stimuli=[dict(target=int(t["target"]), temporal=str(t["temporal"]),
                                     params=[float(v) for v in t.get("params", [])]) for t in d.get("stimuli", [])])
        except (KeyError, TypeError, ValueError) as e:
            _err(L.HB_INVALID_CONFIG, f"config parse error: {e}")

    @classmethod
    def from_toml_str(cls, s: str, base_dir: str = ".") -> "SimConfig":
        try:
            return cls.from_dict(_toml.loads(s), base_dir)
        except _toml.TOMLDecodeError as e:
            _err(L.HB_INVALID_CONFIG, f"TOML parse error: {e}")

    @classmethod
    def from_json_str(cls, s: str, base_dir: str = ".") -> "SimConfig":
        try:
            return cls.from_dict(json.loads(s), base_dir)
        except json.JSONDecodeError as e:
            _err(L.HB_INVALID_CONFIG, f"JSON parse error: {e}")

    @classmethod
    def from_file(cls, path: str) -> "SimConfig":
        try:
            txt = open(path).read()
        except OSError as e:
            _err(L.HB_IO_ERROR, str(e))
        base = os.path.dirname(os.path.abspath(path))
        return cls.from_json_str(txt, base) if path.endswith(".json") else cls.from_toml_str(txt, base)

    def n_steps(self) -> int:
        """(sim_length / dt) as usize: f64 truncation (cli.rs:901)."""
        return int(self.sim_length / self.dt)

    def validate(self) -> None:
        """SimConfig::validate, config.rs:286-428."""
        if self.dt <= 0:
            _err(L.HB_INVALID_CONFIG, "dt must be positive")
        if self.sim_length <= 0:
            _err(L.HB_INVALID_CONFIG, "sim_length must be positive")
        if not self.subnetworks:
            _err(L.HB_INVALID_CONFIG, "At least one subnetwork required")
        for i, s in enumerate(self.subnetworks):
            if s.model not in L.MODEL_NAMES:
                _err(L.HB_INVALID_CONFIG, f"Unknown model '{s.model}' in subnetwork {i}")
            nv, _, npar = L.MODEL_INFO[L.MODEL_NAMES[s.model]]
            if s.nmodes < 1:
                _err(L.HB_INVALID_CONFIG, f"Subnetwork {i}: nmodes must be >= 1")
            if len(s.params) != npar:
                _err(L.HB_INVALID_CONFIG, f"Subnetwork {i}: model '{s.model}' expects {npar} params, got {len(s.params)}")
            if isinstance(s.initial_state, list) and len(s.initial_state) != nv * s.nnodes * s.nmodes:
                _err(L.HB_INVALID_CONFIG, f"Subnetwork {i}: initial_state has {len(s.initial_state)} values, "
                                          f"expected {nv * s.nnodes * s.nmodes}")
        n_subs = len(self.subnetworks)
        for i, p in enumerate(self.projections):
            if p.src >= n_subs or p.tgt >= n_subs:
                _err(L.HB_INVALID_CONFIG, f"Projection {i}: src/tgt out of range")
            ns, nt = self.subnetworks[p.src].nnodes, self.subnetworks[p.tgt].nnodes
            w = p.weights
            if isinstance(w, dict):
                if len(w["indptr"]) != nt + 1:
                    _err(L.HB_INVALID_CONFIG, f"Projection {i}: CSR indptr length {len(w['indptr'])} != tgt_nnodes + 1 = {nt + 1}")
                if len(w["data"]) != len(w["indices"]):
                    _err(L.HB_INVALID_CONFIG, f"Projection {i}: CSR data.len() != indices.len()")
            elif isinstance(w, list):
                if len(w) != nt or any(len(r) != ns for r in w):
                    _err(L.HB_INVALID_CONFIG, f"Projection {i}: dense weights shape mismatch")
            elif ns != nt:
                _err(L.HB_INVALID_CONFIG, f"Projection {i}: scalar weights require src and tgt nnodes to match")
            md = max(p.delays) if p.delays else 0
            if md + 1 > self.n_steps() + 1:
                _err(L.HB_INVALID_CONFIG, f"Projection {i}: max delay {md} requires horizon >= {md + 1}, "
                                          f"but sim only runs {self.n_steps()} steps")

    def resolve_delays(self) -> None:
        """ceil(tract_length / speed / dt), config.rs:436-452."""
        if self.speed <= 0 and any(p.tract_lengths for p in self.projections):
            _err(L.HB_INVALID_CONFIG, "speed must be positive when tract_lengths are provided")
        for p in self.projections:
            if not p.delays and p.tract_lengths:
                dt = np.float32(self.dt)
                p.delays = [int(max(math.ceil(np.float32(l) / np.float32(self.speed) / dt), 0)) for l in p.tract_lengths]


@dataclass
class SweepRange:
    start: float
    step: float
    end: float


@dataclass
class SweepConfig:
    """config.rs:457-503."""
    parameter_name: str
    values: Optional[list] = None
    range: Optional[SweepRange] = None

    @classmethod
    def from_file(cls, path: str) -> "SweepConfig":
        d = _toml.loads(open(path).read())
        r = d.get("range")
        return cls(parameter_name=d["parameter_name"], values=d.get("values"),
                   range=SweepRange(**r) if r else None)

    def generate_values(self) -> list:
        if self.values is not None:
            return [float(np.float32(v)) for v in self.values]
        if self.range is not None:
            st, sp, en = (np.float32(self.range.start), np.float32(self.range.step), np.float32(self.range.end))
            if sp <= 0:
                return []
            n = int(math.ceil(np.float32((en - st) / sp))) + 1
            return [float(np.float32(st + np.float32(i) * sp)) for i in range(n)]
        return []


def parse_sweep_param(name: str, cfg: SimConfig):
    """'subnetworks[i].params[j]' -> (i, j), cli.rs:976-999."""
    if name in ("dt", "nsig"):
        _err(L.HB_INVALID_CONFIG, f"Batch sweep does not support sweeping {name} (model parameter sweeps only)")
    if name.startswith("subnetworks[") and "].params[" in name:
        sub = int(name[name.index("[") + 1:name.index("]")])
        par = int(name[name.rindex("[") + 1:name.rindex("]")])
        if sub >= len(cfg.subnetworks):
            _err(L.HB_INVALID_CONFIG, f"subnetwork index {sub} out of range")
        return sub, par
    _err(L.HB_INVALID_CONFIG, f"unsupported sweep parameter_name: {name}. Use format: subnetworks[N].params[M]")


@dataclass
class SweepParam:
    sub_idx: int
    param_idx: int


@dataclass
class SweepPoint:
    sub_idx: int
    param_idx: int
    values: list
    kind: int = 0          # 0: model parameter, 1: coupling parameter (extension)


@dataclass
class SweepSpec:
    points: list

    def n_sweep(self) -> int:
        n = 1
        for p in self.points:
            n *= len(p.values)
        return n

    def cartesian_product(self) -> list:
        """Last dimension fastest, engine.rs:69-90."""
        out = []
        idx = [0] * len(self.points)
        for _ in range(self.n_sweep()):
            out.append([(p.sub_idx, p.param_idx, p.values[idx[i]]) for i, p in enumerate(self.points)])
            for i in range(len(self.points) - 1, -1, -1):
                idx[i] += 1
                if idx[i] < len(self.points[i].values):
                    break
                idx[i] = 0
        return out


@dataclass
class BatchSweepResult:
    """engine.rs:15-31 (tavg/final per subnetwork, flat [n_sweep, nnodes, nvar]).
    bold: one [series][rows][nnodes] array, the whole series since the engine
    started resampled to TR; series = n_sweep (a monitor per sweep point) or
    1 (bold_batch_mean: the reference's monitor of the sweep mean,
    engine.rs:785-801). On the multi-subnetwork path the per-point series
    cover the run only; the batch-mean monitor carries its series across runs
    as the reference's does."""
    param_values: list
    tavg: list
    final_states: list
    trajectories: Optional[list]
    bold: Optional[list]
    elapsed_ms: float
    n_sweep: int
    kernel_ms: float = 0.0


# ---------------------------------------------------------------- engine ----

def _load_initial_state(sub: SubnetworkConfig, nv: int, base_dir: str) -> np.ndarray:
    """Node-major [nnodes, nvar] initial state. Inline values are node-major in
    the batch engine (engine.rs:161-170); a 2-D .npy [nvar, nnodes] is
    transposed (engine.rs:177-196)."""
    s = sub.initial_state
    if isinstance(s, str):
        path = s if os.path.isabs(s) else os.path.join(base_dir, s)
        try:
            a = np.load(path, allow_pickle=False)
        except OSError as e:
            _err(L.HB_IO_ERROR, str(e))
        if a.ndim == 2:
            return np.ascontiguousarray(a.astype(np.float32).T)
        if a.ndim == 3 and a.shape[2] == 1:
            return np.ascontiguousarray(a[:, :, 0].astype(np.float32).T)
        _err(L.HB_INVALID_CONFIG, f"Initial state NPY has {a.ndim} dims, expected 2 or 3")
    return np.asarray(s, np.float32).reshape(sub.nnodes, nv)


def _needs_multi(cfg: SimConfig) -> bool:
    """Several subnetworks, several projections, a projection between two
    subnetworks or a cvar_map other than 0:0 take the multi-subnetwork path."""
    return (len(cfg.subnetworks) != 1 or len(cfg.projections) > 1
            or any(p.src != p.tgt or p.cvar_map != "0:0" for p in cfg.projections)
            or L.MODEL_NAMES.get(cfg.subnetworks[0].model, 0) in L.MULTI_ONLY)


def _multi_from_config(cfg: SimConfig, plain: bool = False):
    """SimConfig -> (Subnet list, Link list) for hyburn_amd.multi (engine.rs:145-415).
    plain: a CSR projection with one delay per edge keeps its per-edge delays
    (the HybridEngine's path, runtime.rs:125-205); otherwise CSR is densified."""
    from .multi import Link, Subnet, parse_cvar_map
    subs = []
    for i, s in enumerate(cfg.subnetworks):
        if s.model not in L.MODEL_NAMES:
            _err(L.HB_INVALID_CONFIG, f"Unknown model '{s.model}' in subnetwork {i}")
        if s.nmodes != 1:
            _err(L.HB_INVALID_CONFIG, "nmodes > 1 is not supported by this backend")
        m = L.MODEL_NAMES[s.model]
        subs.append(Subnet(m, s.nnodes, list(s.params), _load_initial_state(s, L.MODEL_INFO[m][0], cfg.base_dir)))
    links = []
    for k, p in enumerate(cfg.projections):
        if not (0 <= p.src < len(subs) and 0 <= p.tgt < len(subs)):
            _err(L.HB_INVALID_CONFIG, f"projection {k}: src/tgt out of range")
        if p.coupling_fn not in L.COUPLING_NAMES:
            _err(L.HB_INVALID_CONFIG, f"Unknown coupling function: {p.coupling_fn}")
        ns, nt = subs[p.src].n_nodes, subs[p.tgt].n_nodes
        w = p.weights
        if isinstance(w, dict) and plain and len(p.delays) == len(w["data"]) and len(p.delays) > 0:
            links.append(Link(p.src, p.tgt, (np.asarray(w["indptr"]), np.asarray(w["indices"]), np.asarray(w["data"])),
                              L.COUPLING_NAMES[p.coupling_fn], list(p.coupling_params), parse_cvar_map(p.cvar_map),
                              0, edge_delays=list(p.delays)))
            continue
        if isinstance(w, dict):   # CSR by target row, densified (engine.rs:298-311)
            W = np.zeros((nt, ns), np.float32)
            ip, ix, dat = w["indptr"], w["indices"], w["data"]
            for r in range(nt):
                for e in range(ip[r], ip[r + 1]):
                    W[r, ix[e]] = dat[e]
            w = W
        elif isinstance(w, (list, np.ndarray)):
            w = np.asarray(w, np.float32).reshape(nt, ns)
        else:
            w = float(w)
        links.append(Link(p.src, p.tgt, w, L.COUPLING_NAMES[p.coupling_fn], list(p.coupling_params),
                          parse_cvar_map(p.cvar_map), int(p.delays[0]) if p.delays else 0))
    return subs, links


def multi_checkpoint_states(c: dict, subs, links, max_delay: int, n_sweep: int):
    """Check a read checkpoint against a multi-subnetwork configuration
    (HybridEngine::resume's shape checks, runtime.rs:693-725) and return its
    states and history for n_sweep points; a file of one point continues
    every point."""
    from .multi import horizon
    if len(c["shapes"]) != len(subs):
        _err(L.HB_INVALID_STATE, f"Checkpoint has {len(c['shapes'])} subnetworks, expected {len(subs)}")
    for i, ((nv, nn, nm), S) in enumerate(zip(c["shapes"], subs)):
        want = (L.MODEL_INFO[S.model][0], S.n_nodes, 1)
        if (nv, nn, nm) != want:
            _err(L.HB_INVALID_STATE, f"Checkpoint shape mismatch for subnetwork {i}: expected "
                                     f"(nvar={want[0]},nnodes={want[1]},nmodes=1), got ({nv},{nn},{nm})")
    if c.get("hist_single") is not None:
        _err(L.HB_INVALID_STATE, "this checkpoint carries a single-network engine's history (version 4)")
    states, hist = c["states"], c["hist"]
    H = horizon(links, max_delay)
    n_tot = sum(S.n_nodes for S in subs)
    if hist is not None and (hist.shape[1] != n_tot or hist.shape[2] != H):
        _err(L.HB_INVALID_STATE, f"Checkpoint history has shape {tuple(hist.shape)}, expected "
                                 f"(n_sweep, {n_tot}, {H}, 2)")
    if c["n_sweep"] != n_sweep:
        if c["n_sweep"] != 1:
            _err(L.HB_INVALID_STATE, f"Checkpoint holds {c['n_sweep']} sweep points, expected {n_sweep}")
        # one saved point continues every sweep point: states and history alike
        states = [np.repeat(a, n_sweep, axis=0) for a in states]
        if hist is not None:
            hist = np.repeat(hist, n_sweep, axis=0)
    return states, hist


class BatchHybridEngine:
    """BatchHybridEngine<B> with B = the HIP backend. Configurations with
    several subnetworks (or projections between subnetworks) run through
    hb_run_multi; set hybrid_integrator = True for Heun on G2DO and the
    configured rule elsewhere (engine.rs:698-702)."""

    def __init__(self):
        self._h = C.c_void_p()
        self._keep: list = []
        self._multi = None
        self._max_delay = 0
        self._nsig, self._noise_seed = 0.0, 0
        self._step, self._state, self._hist = 0, None, None   # carried across run_sweep calls
        self._bold = None
        self.hybrid_integrator = False
        # one BOLD monitor fed the sweep mean of var0 (the reference's batch
        # monitor, engine.rs:785-801) instead of one per sweep point
        self.bold_batch_mean = False
        # that monitor on the multi-subnetwork path, carried across runs:
        # Balloon-Windkessel state, native samples and their times
        self._bmon_state, self._bmon_native, self._bmon_times = None, [], []
        self.step = 0

    @classmethod
    def from_config(cls, base_config: SimConfig, n_sweep: int, device: int = 0, *,
                    engine: int = L.ENGINE_BATCH, delay_mode: int = -1, noise_seed: int = 0,
                    noise_shared: bool = True, bold_batch_mean: bool = False) -> "BatchHybridEngine":
        self = cls._build(base_config, n_sweep, device, engine, delay_mode, noise_seed, noise_shared)
        self.bold_batch_mean = bool(bold_batch_mean)
        return self

    @classmethod
    def _build(cls, base_config, n_sweep, device, engine, delay_mode, noise_seed, noise_shared):
        cfg = base_config
        if _needs_multi(cfg):
            # the multi-subnetwork branch runs the batch engine's step; the plain
            # engine's multi-subnetwork step is run_multi(plain=True), which the
            # run command calls, and the single-network delay/noise options do
            # not apply to it
            if engine != L.ENGINE_BATCH:
                _err(L.HB_INVALID_CONFIG, "plain-engine runs of this configuration go through "
                                          "hyburn_amd.multi.run_multi(plain=True), not from_config")
            if delay_mode != -1 or not noise_shared:
                _err(L.HB_INVALID_CONFIG, "delay_mode and noise_shared apply to single-network "
                                          "configurations only")
            self = cls()
            if cfg.integrator not in L.INTEGRATOR_NAMES:
                _err(L.HB_INVALID_CONFIG, f"unknown integrator {cfg.integrator}")
            subs, links = _multi_from_config(cfg)
            bold = [m for m in cfg.monitors if m.monitor_type.lower() == "bold"]
            if bold:   # engine.rs:418-447: subnetwork 0, windows of bold_period steps
                m = bold[0]
                self._bold = (int(m.bold_period if m.bold_period is not None
                                  else max(1, round((m.period or 2000.0) / cfg.dt))),
                              float(m.tr if m.tr is not None else 2.0) * 1000.0)
            self._multi = (subs, links, L.INTEGRATOR_NAMES[cfg.integrator], device)
            self._max_delay = max([max(p.delays) for p in cfg.projections if p.delays] or [0])
            self._nsig, self._noise_seed = cfg.nsig, noise_seed
            self.n_sweep, self.dt = n_sweep, cfg.dt
            return self
        sub = cfg.subnetworks[0]
        if sub.model not in L.MODEL_NAMES:
            _err(L.HB_INVALID_CONFIG, f"Unknown model '{sub.model}'")
        if sub.nmodes != 1:
            _err(L.HB_INVALID_CONFIG, "nmodes > 1 is not supported by this backend")
        model = L.MODEL_NAMES[sub.model]
        nv = L.MODEL_INFO[model][0]
        self = cls()
        k = self._keep
        self.model, self.nnodes, self.nvar, self.n_sweep = model, sub.nnodes, nv, n_sweep
        self.dt = cfg.dt
        params = np.ascontiguousarray(sub.params, np.float32)
        x0 = np.ascontiguousarray(_load_initial_state(sub, nv, cfg.base_dir))
        k += [params, x0]
        S = L.HbSubnetwork(model, sub.nnodes, len(params), params.ctypes.data_as(L.fp), x0.ctypes.data_as(L.fp))
        k.append(S)
        c = L.HbSimConfig()
        c.sim_length, c.dt = cfg.sim_length, cfg.dt
        if cfg.integrator not in L.INTEGRATOR_NAMES:
            _err(L.HB_INVALID_CONFIG, f"unknown integrator {cfg.integrator}")
        c.integrator = L.INTEGRATOR_NAMES[cfg.integrator]
        nsig = np.atleast_1d(np.asarray(cfg.nsig, np.float32))
        k.append(nsig)
        c.n_nsig, c.nsig = len(nsig), nsig.ctypes.data_as(L.fp)
        c.subnetworks, c.n_subnetworks = C.pointer(S), 1
        if cfg.projections:
            if len(cfg.projections) > 1:
                _err(L.HB_INVALID_CONFIG, "at most one (self) projection is supported by this backend")
            p = cfg.projections[0]
            if p.coupling_fn not in L.COUPLING_NAMES:
                _err(L.HB_INVALID_CONFIG, f"Unknown coupling function: {p.coupling_fn}")
            Pj = L.HbProjection()
            w = p.weights
            if isinstance(w, dict):
                ip = np.ascontiguousarray(w["indptr"], np.int32)
                ix = np.ascontiguousarray(w["indices"], np.int32)
                dat = np.ascontiguousarray(w["data"], np.float32)
                k += [ip, ix, dat]
                Pj.wkind, Pj.indptr, Pj.indices, Pj.data, Pj.nnz = (L.W_CSR, ip.ctypes.data_as(L.ip),
                                                                    ix.ctypes.data_as(L.ip), dat.ctypes.data_as(L.fp), len(dat))
            elif isinstance(w, (list, np.ndarray)):
                dn = np.ascontiguousarray(w, np.float32)
                k.append(dn)
                Pj.wkind, Pj.dense = L.W_DENSE, dn.ctypes.data_as(L.fp)
            else:
                Pj.wkind, Pj.scalar_w = L.W_SCALAR_MEAN, float(w)
            dl = np.ascontiguousarray(p.delays, np.int32)
            k.append(dl)
            Pj.delays, Pj.n_delays = (dl.ctypes.data_as(L.ip) if len(dl) else None), len(dl)
            Pj.cfn = L.COUPLING_NAMES[p.coupling_fn]
            cp = list(p.coupling_params)[:8]
            Pj.n_cparams = len(cp)
            for i, v in enumerate(cp):
                Pj.cparams[i] = v
            k.append(Pj)
            c.projections, c.n_projections = C.pointer(Pj), 1
        bold = [m for m in cfg.monitors if m.monitor_type.lower() == "bold"]
        if bold:
            m = bold[0]
            bp = m.bold_period if m.bold_period is not None else max(1, round((m.period or 2000.0) / cfg.dt))
            c.bold_period, c.bold_tr_s = int(bp), float(m.tr if m.tr is not None else 2.0)
        c.engine, c.delay_mode, c.noise_seed, c.noise_shared = engine, delay_mode, noise_seed, int(noise_shared)
        self._cfg = c
        self._pycfg, self._device = cfg, device
        L.check(L.lib().hb_engine_create(C.byref(c), n_sweep, device, C.byref(self._h)))
        return self

    def _result(self, r: L.HbResult, values, traj: bool) -> BatchSweepResult:
        n = r.n_sweep * r.n_nodes * r.nvar
        tavg = np.ctypeslib.as_array(r.tavg, (n,)).copy()
        fin = np.ctypeslib.as_array(r.final_states, (n,)).copy()
        trj = np.ctypeslib.as_array(r.trajectories, (n * r.n_steps,)).copy() if traj and r.n_steps else None
        bold = None
        if r.bold and r.n_bold > 0:
            # one series per sweep point, or one fed the batch mean (bold_batch_mean)
            bold = [np.ctypeslib.as_array(r.bold, (r.bold_series * r.n_bold * r.n_nodes,)).copy()
                    .reshape(r.bold_series, r.n_bold, r.n_nodes)]
        self.step = int(r.step)
        out = BatchSweepResult(param_values=list(values), tavg=[tavg], final_states=[fin],
                               trajectories=[trj] if trj is not None else None, bold=bold,
                               elapsed_ms=r.elapsed_ms, n_sweep=r.n_sweep, kernel_ms=r.kernel_ms)
        L.lib().hb_result_free(C.byref(r))
        return out

    def _run(self, param: SweepParam, values, n_steps: int, traj: bool, exact: bool,
             restart: bool = False) -> BatchSweepResult:
        v = np.ascontiguousarray(values, np.float32)
        r = L.HbResult()
        flags = ((L.FLAG_TRAJECTORY if traj else 0) | (L.FLAG_EXACT if exact else 0)
                 | (L.FLAG_RESTART if restart else 0) | (L.FLAG_BOLD_BATCH_MEAN if self.bold_batch_mean else 0))
        L.check(L.lib().hb_run_sweep(self._h, param.sub_idx, param.param_idx, v.ctypes.data_as(L.fp), len(v),
                                     int(n_steps), flags, C.byref(r)))
        return self._result(r, v.tolist(), traj)

    def _run_multi(self, param: SweepParam, values, n_steps: int, traj: bool = False) -> BatchSweepResult:
        from .multi import run_multi
        subs, links, integ, device = self._multi
        v = np.ascontiguousarray(values, np.float32)
        if len(v) != self.n_sweep:
            _err(L.HB_INVALID_CONFIG, "number of sweep values must equal n_sweep")
        if not 0 <= param.sub_idx < len(subs):
            _err(L.HB_INVALID_CONFIG, "subnetwork index out of range")
        if self._state is not None and len(self._state[0]) != self.n_sweep:
            _err(L.HB_INVALID_STATE, f"the engine holds {len(self._state[0])} sweep points (left by run_sweep_multi); "
                                     f"this run has {self.n_sweep}")
        t0 = time.perf_counter()
        # like the reference's engine, a run continues from the previous run's
        # states, step counter and history (engine.rs:496-540)
        r = run_multi(subs, links, dt=self.dt, n_steps=int(n_steps), integrator=integ, nsig=self._nsig,
                      noise_seed=self._noise_seed + self._step,
                      hybrid_integrator=self.hybrid_integrator, n_sweep=self.n_sweep,
                      sweep=(param.sub_idx, param.param_idx, v), max_delay=int(self._max_delay),
                      step0=self._step, state=self._state, hist=self._hist, record_trajectory=traj,
                      bold_period=self._bold[0] if self._bold else 0,
                      bold_tr_ms=self._bold[1] if self._bold else 0.0, device=device,
                      bold_mean=self.bold_batch_mean, bold_state=self._bmon_state)
        ms = (time.perf_counter() - t0) * 1e3
        self._step, self._state, self._hist = r.step_end, r.final_states, r.hist
        bold = r.bold
        if r.bold is not None and self.bold_batch_mean:
            bold = self._bold_mean_series(r, int(n_steps))
        return BatchSweepResult(param_values=v.tolist(), tavg=[t.ravel() for t in r.tavg],
                                final_states=[f.ravel() for f in r.final_states],
                                trajectories=[t.ravel() for t in r.trajectories] if traj else None,
                                bold=[bold] if bold is not None else None,
                                elapsed_ms=ms, n_sweep=self.n_sweep, kernel_ms=r.kernel_ms)

    def _bold_mean_series(self, r, n_steps: int) -> np.ndarray:
        """The batch monitor's whole series since the engine started, [1][rows][nnodes]:
        this run's native samples appended at the times BoldMonitor gives them
        (full windows end at len * dtb + dtb, the flushed partial window at
        len * dtb + dt * rem; bold_monitor.rs:100, 214), resampled to TR by
        hb_bold_downsample (bold_monitor.rs:123-203)."""
        P, tr_ms = self._bold
        dtb = float(self.dt) * P
        n_full, rem = divmod(n_steps, P)
        for k in range(r.bold.shape[0]):
            ln = len(self._bmon_times)
            self._bmon_times.append(ln * dtb + dtb if k < n_full else ln * dtb + float(self.dt) * rem)
        self._bmon_native.append(r.bold)
        self._bmon_state = r.bold_state
        nat = np.ascontiguousarray(np.concatenate(self._bmon_native), np.float32)
        tm = np.ascontiguousarray(self._bmon_times, np.float64)
        dp = C.POINTER(C.c_double)
        N = nat.shape[1]
        rows = int(L.lib().hb_bold_downsample(tm.ctypes.data_as(dp), len(tm), nat.ctypes.data_as(L.fp), N, dtb,
                                              tr_ms, None, 0))
        out = np.zeros((max(rows, 1), N), np.float32)
        got = int(L.lib().hb_bold_downsample(tm.ctypes.data_as(dp), len(tm), nat.ctypes.data_as(L.fp), N, dtb,
                                             tr_ms, out.ctypes.data_as(L.fp), out.shape[0]))
        if got != rows:
            _err(L.HB_BACKEND_ERROR, "BOLD resampling returned an unexpected row count")
        return out[:rows][None]

    def _hybrid_changes_rule(self) -> bool:
        # hybrid_integrator: Heun for G2DO, else the configured integrator with
        # a configured Heun read as Euler (engine.rs:698-702)
        if not self.hybrid_integrator or self._multi is not None:
            return False
        heun = self._cfg.integrator == L.HEUN
        return (self.model == L.G2DO) != heun

    def run_sweep(self, param: SweepParam, param_values, n_steps: int, exact: bool = False,
                  restart: bool = False) -> BatchSweepResult:
        """Run n_steps from where the previous run stopped (engine.rs:496-540);
        restart=True starts from the configured initial state at step 0."""
        if self._multi is None and self._hybrid_changes_rule():
            cfg = self._pycfg
            if any(m.monitor_type.lower() == "bold" for m in cfg.monitors):
                _err(L.HB_INVALID_CONFIG, "hybrid_integrator needs a run without BOLD here")
            subs, links = _multi_from_config(cfg)
            self._multi = (subs, links, self._cfg.integrator, self._device)
            self._max_delay = max([max(p.delays) for p in cfg.projections if p.delays] or [0])
            self._nsig = cfg.nsig
        if self._multi is not None:
            return self._run_multi(param, param_values, n_steps)
        return self._run(param, param_values, n_steps, False, exact, restart)

    def run_sweep_with_trajectory(self, param: SweepParam, param_values, n_steps: int,
                                  exact: bool = False, restart: bool = False) -> BatchSweepResult:
        if self._multi is not None:
            return self._run_multi(param, param_values, n_steps, traj=True)
        return self._run(param, param_values, n_steps, True, exact, restart)

    def run_sweep_multi(self, spec: SweepSpec, n_steps: int, exact: bool = False,
                        trajectory: bool = False) -> BatchSweepResult:
        """The Cartesian product of spec's dimensions (engine.rs:845-851, 862-1348):
        a fresh run of n_sweep() points from sweep point 0 of the current
        state at step 0, with empty histories and no monitors, whose states,
        step and histories the engine then keeps. param_values lists every
        point's values, dimension by dimension."""
        if self._multi is not None:
            return self._run_sweep_multi_subnets(spec, n_steps, trajectory)
        arr = (L.HbSweepPoint * len(spec.points))()
        keep = []
        for i, p in enumerate(spec.points):
            v = np.ascontiguousarray(p.values, np.float32)
            keep.append(v)
            arr[i] = L.HbSweepPoint(p.kind, p.sub_idx, p.param_idx, v.ctypes.data_as(L.fp), len(v))
        r = L.HbResult()
        flags = (L.FLAG_TRAJECTORY if trajectory else 0) | (L.FLAG_EXACT if exact else 0)
        L.check(L.lib().hb_run_sweep_multi(self._h, arr, len(spec.points), int(n_steps), flags, C.byref(r)))
        return self._result(r, [v for pt in spec.cartesian_product() for (_, _, v) in pt], trajectory)

    def _run_sweep_multi_subnets(self, spec: SweepSpec, n_steps: int, traj: bool) -> BatchSweepResult:
        from .multi import Subnet, run_multi
        subs, links, integ, device = self._multi
        if any(p.kind != 0 for p in spec.points):
            _err(L.HB_INVALID_CONFIG, "coupling-parameter dimensions are swept on single-network configurations only")
        n = spec.n_sweep()
        if n < 1:
            _err(L.HB_INVALID_CONFIG, "empty sweep spec")
        t0 = time.perf_counter()
        # the fresh engine starts from sweep point 0 of the current states (engine.rs:886-897)
        start = [Subnet(S.model, S.n_nodes, S.params,
                        np.asarray(self._state[k][0], np.float32) if self._state is not None else S.x0)
                 for k, S in enumerate(subs)]
        dims = [(p.sub_idx, p.param_idx, list(p.values)) for p in spec.points]
        r = run_multi(start, links, dt=self.dt, n_steps=int(n_steps), integrator=integ, nsig=self._nsig,
                      noise_seed=self._noise_seed, hybrid_integrator=self.hybrid_integrator, n_sweep=n,
                      sweep_spec=dims, max_delay=int(self._max_delay), record_trajectory=traj, device=device)
        ms = (time.perf_counter() - t0) * 1e3
        # engine.rs:1326-1331: the caller's engine takes over states, step and
        # histories, and the local engine's monitors, of which it has none
        # (engine.rs:899-905): later runs record no BOLD
        self._state, self._hist, self._step = r.final_states, r.hist, int(n_steps)
        self._bold, self._bmon_state, self._bmon_native, self._bmon_times = None, None, [], []
        return BatchSweepResult(param_values=[v for pt in spec.cartesian_product() for (_, _, v) in pt],
                                tavg=[t.ravel() for t in r.tavg], final_states=[f.ravel() for f in r.final_states],
                                trajectories=[t.ravel() for t in r.trajectories] if traj else None, bold=None,
                                elapsed_ms=ms, n_sweep=n, kernel_ms=r.kernel_ms)

    def checkpoint(self, path: str, reference_format: Optional[bool] = None) -> None:
        """Save the engine state (HybridEngine::checkpoint, runtime.rs:519-561).

        Single-network engines write the reference's version 2 file when that
        loses nothing (one sweep point, no delay history to carry) and the
        version 4 batch file otherwise; reference_format True forces version
        2 (one sweep point; a delay history is dropped, as the reference
        keeps none), False forces version 4. Multi-subnetwork engines write
        version 3 (states and history of every sweep point)."""
        from . import checkpoint as ck
        if self._multi is not None:
            if reference_format:
                _err(L.HB_INVALID_STATE, "a multi-subnetwork engine writes the batch (version 3) checkpoint")
            subs = self._multi[0]
            shapes = [(L.MODEL_INFO[S.model][0], S.n_nodes) for S in subs]
            states = self._state or [np.broadcast_to(np.asarray(S.x0, np.float32), (self.n_sweep,) + np.shape(S.x0))
                                     for S in subs]
            nsig = np.atleast_1d(np.asarray(self._nsig, np.float32))
            ck.write_batch(path, step=self._step, dt=self.dt, integrator=self._multi[2], nsig=nsig, shapes=shapes,
                           states=states, hist=self._hist)
            return
        st, nsig, states, hist = self._get_state()
        carried = st.n_state > 0 and st.hist_floats > 0
        v2 = states.shape[0] == 1 and not carried if reference_format is None else bool(reference_format)
        if v2:
            if states.shape[0] != 1:
                _err(L.HB_INVALID_STATE, f"the reference's checkpoint holds one simulation; this engine holds "
                                         f"{states.shape[0]} sweep points")
            ck.write_reference(path, step=st.step, dt=st.dt, integrator=st.integrator, nsig=nsig, nvar=self.nvar,
                               nnodes=self.nnodes, state=states[0])
        else:
            ck.write_single(path, step=st.step, dt=st.dt, integrator=st.integrator, nsig=nsig, nvar=self.nvar,
                            nnodes=self.nnodes, states=states, hist=hist if carried else None)

    def _get_state(self):
        st = L.HbEngineState()
        L.check(L.lib().hb_engine_get_state(self._h, C.byref(st), None, None, None))
        states = np.zeros((max(st.n_state, 1), self.nnodes, self.nvar), np.float32)
        nsig = np.zeros(self.nvar, np.float32)
        hist = np.zeros((st.n_state, st.hist_floats), np.float32) if st.n_state and st.hist_floats else None
        L.check(L.lib().hb_engine_get_state(self._h, C.byref(st), nsig.ctypes.data_as(L.fp),
                                            states.ctypes.data_as(L.fp),
                                            hist.ctypes.data_as(L.fp) if hist is not None else None))
        return st, nsig, states, hist

    def resume(self, path: str) -> None:
        """Restore step, dt, integrator, nsig and the states from a HYBURNCK
        file (HybridEngine::resume, runtime.rs:576-742); the shapes must
        match. A file of one sweep point continues every sweep point; one
        without history (the reference's versions 1 and 2) continues from a
        zero delay history, which is what a freshly built reference engine
        resumes into."""
        from . import checkpoint as ck
        c = ck.read(path)
        if self._multi is not None:
            self._resume_multi(c)
            return
        if len(c["shapes"]) != 1:
            _err(L.HB_INVALID_STATE, f"Checkpoint has {len(c['shapes'])} subnetworks, expected 1")
        nv, nn, nm = c["shapes"][0]
        if (nv, nn, nm) != (self.nvar, self.nnodes, 1):
            _err(L.HB_INVALID_STATE, f"Checkpoint shape mismatch for subnetwork 0: expected "
                                     f"(nvar={self.nvar},nnodes={self.nnodes},nmodes=1), got ({nv},{nn},{nm})")
        if c["hist"] is not None:
            _err(L.HB_INVALID_STATE, "this checkpoint carries a multi-subnetwork history; "
                                     "a single-network engine keeps its own layout (version 4)")
        states = np.ascontiguousarray(c["states"][0], np.float32)
        hist = c["hist_single"]
        n = states.shape[0]
        if n != 1 and n != self.n_sweep:
            _err(L.HB_INVALID_STATE, f"Checkpoint holds {n} sweep points, expected 1 or {self.n_sweep}")
        st = L.HbEngineState(int(c["step"]), float(c["dt"]), int(c["integrator"]), n, self.nnodes, self.nvar,
                             0 if hist is None else hist.shape[1])
        nsig = np.ascontiguousarray(c["nsig"], np.float32)
        if hist is not None:
            hist = np.ascontiguousarray(hist, np.float32)
        L.check(L.lib().hb_engine_set_state(self._h, C.byref(st), nsig.ctypes.data_as(L.fp), len(nsig),
                                            states.ctypes.data_as(L.fp),
                                            hist.ctypes.data_as(L.fp) if hist is not None else None))
        self.step, self.dt = int(c["step"]), float(c["dt"])

    def _resume_multi(self, c: dict) -> None:
        subs, links, _, device = self._multi
        states, hist = multi_checkpoint_states(c, subs, links, int(self._max_delay), self.n_sweep)
        self._step, self.dt, self._nsig = int(c["step"]), float(c["dt"]), c["nsig"]
        self._multi = (subs, links, c["integrator"], device)
        self._state, self._hist = states, hist

    def close(self):
        if self._h:
            L.lib().hb_engine_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- NPY -------

def write_npy_f32(path: str, data, shape) -> None:
    """Byte-compatible with io.rs:88-125: v1.0 header, '<f4', padded so that
    magic + header + newline is a multiple of 64."""
    shape = list(shape)
    shape_str = f"({shape[0]},)" if len(shape) == 1 else "(" + ", ".join(str(s) for s in shape) + ")"
    header = "{'descr': '<f4', 'fortran_order': False, 'shape': " + shape_str + " }"
    rem = (len(header) + 10 + 1) % 64
    header += " " * ((64 - rem) if rem else 0) + "\n"
    arr = np.ascontiguousarray(data, dtype="<f4").ravel()
    with open(path, "wb") as f:
        f.write(b"\x93NUMPY")
        f.write(bytes([1, 0]))
        f.write(len(header).to_bytes(2, "little"))
        f.write(header.encode("latin1"))
        f.write(arr.tobytes())
