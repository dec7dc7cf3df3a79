"""Instance-level API: one fully specified simulation per `Instance`, run as a
ragged batch by the HIP planner (hb_plan_* in include/hyburn_hip.h).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib as L


def _f32(a) -> Optional[np.ndarray]:
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _i32(a) -> Optional[np.ndarray]:
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _fptr(a):
    return None if a is None else a.ctypes.data_as(L.fp)


def _iptr(a):
    return None if a is None else a.ctypes.data_as(L.ip)


@dataclass
class Instance:
    """One simulation: a single network with a self-projection."""

    model: int
    n_nodes: int
    n_steps: int
    dt: float                       # ms; held as f32 by the engine
    params: np.ndarray
    x0: np.ndarray                  # [n_nodes, nvar] node-major
    integrator: int = L.HEUN
    engine: int = L.ENGINE_BATCH
    dt_ms: Optional[float] = None   # f64 config dt for BOLD window times (default: dt)
    sweep_form: int = 0
    sweep_val: float = 0.0
    node_param_idx: int = -1
    node_param: Optional[np.ndarray] = None
    nsig: Optional[np.ndarray] = None
    noise: Optional[np.ndarray] = None   # raw N(0,1) [n_steps, n_nodes, nvar]
    noise_seed: int = 0
    noise_stream: int = 0
    wkind: int = L.W_NONE
    cfn: int = L.CF_LINEAR
    cparams: Sequence[float] = ()
    scalar_w: float = 0.0
    dense: Optional[np.ndarray] = None   # [n_nodes, n_nodes], row = target
    indptr: Optional[np.ndarray] = None
    indices: Optional[np.ndarray] = None
    weights: Optional[np.ndarray] = None
    edge_delays: Optional[np.ndarray] = None
    delay_mode: int = L.DELAY_PER_EDGE
    single_delay: int = 0
    max_delay_hint: int = 0
    bold_period: int = 0
    bold_tr_ms: float = 0.0
    record_traj: bool = False
    record_bold: bool = False
    # continuation of an earlier run (hb_instance.step0 / history / bold_state):
    # steps already integrated, the carried pre() history (float32 array of
    # hb_history_floats() values, read when step0 > 0 and overwritten with the
    # history at the run's end) and the BOLD monitor state [n_nodes][4] (in/out)
    step0: int = 0
    history: Optional[np.ndarray] = None
    bold_state: Optional[np.ndarray] = None

    @property
    def nvar(self) -> int:
        return L.MODEL_INFO[self.model][0]

    @property
    def ncvar(self) -> int:
        return L.MODEL_INFO[self.model][1]

    @property
    def nnz(self) -> int:
        return 0 if self.indices is None else int(len(self.indices))

    def max_delay(self) -> int:
        """Largest delay the ring holds (the reference horizon minus one)."""
        if self.wkind == L.W_NONE:
            return 0
        if self.delay_mode == L.DELAY_BATCH_COMPAT:
            m = max(self.single_delay, self.max_delay_hint)
            if self.edge_delays is not None and len(self.edge_delays):
                m = max(m, int(np.max(self.edge_delays)))
            return m
        if self.wkind == L.W_CSR and self.delay_mode == L.DELAY_PER_EDGE:
            return int(np.max(self.edge_delays)) if self.edge_delays is not None and len(self.edge_delays) else 0
        return self.single_delay

    def bold_rows(self) -> int:
        """BOLD rows produced: native samples, or floor(max_t / TR) after
        resampling (bold_monitor.rs:72-227)."""
        if self.bold_period <= 0:
            return 0
        dt_ms = self.dt if self.dt_ms is None else self.dt_ms
        P, T = self.bold_period, self.n_steps
        dtb = dt_ms * P
        n_full, rem = divmod(T, P)
        n_native = n_full + (1 if rem else 0)
        if self.bold_tr_ms <= 0.0 or abs(self.bold_tr_ms - dtb) < 1e-6:
            return n_native
        if n_native == 0:
            return 0
        max_t = n_full * dtb + dt_ms * rem if rem else (n_full - 1) * dtb + dtb
        return int(math.floor(max_t / self.bold_tr_ms)) if max_t >= self.bold_tr_ms else 0


@dataclass
class Output:
    tavg: np.ndarray
    final_state: np.ndarray
    traj: Optional[np.ndarray] = None
    bold: Optional[np.ndarray] = None


class _Marshal:
    """Keeps numpy buffers alive while the C structs point into them."""

    def __init__(self, insts: Sequence[Instance]):
        self.keep: list = []
        self.outputs: list[Output] = []
        self.bold_n = np.zeros(len(insts), dtype=np.int32)
        self.arr = (L.HbInstance * len(insts))()
        for i, I in enumerate(insts):
            self._one(i, I)

    def _k(self, a):
        if a is not None:
            self.keep.append(a)
        return a

    def _one(self, i: int, I: Instance):
        s = self.arr[i]
        nv = I.nvar
        s.model, s.integrator, s.engine = I.model, I.integrator, I.engine
        s.n_nodes, s.n_steps = I.n_nodes, I.n_steps
        s.dt = float(np.float32(I.dt))
        s.dt_ms = float(I.dt if I.dt_ms is None else I.dt_ms)
        s.params = _fptr(self._k(_f32(I.params)))
        s.sweep_form, s.sweep_val = int(I.sweep_form), float(I.sweep_val)
        s.node_param_idx = int(I.node_param_idx)
        s.node_param = _fptr(self._k(_f32(I.node_param)))
        x0 = self._k(_f32(np.asarray(I.x0).reshape(I.n_nodes, nv)))
        s.x0 = _fptr(x0)
        s.nsig = _fptr(self._k(_f32(I.nsig)))
        s.noise = _fptr(self._k(_f32(I.noise)))
        s.noise_seed, s.noise_stream = int(I.noise_seed), int(I.noise_stream)
        s.wkind, s.cfn = I.wkind, I.cfn
        cp = list(I.cparams)[:8]
        s.n_cparams = len(cp)
        for k, v in enumerate(cp):
            s.cparams[k] = float(v)
        s.scalar_w = float(I.scalar_w)
        s.dense = _fptr(self._k(_f32(I.dense)))
        s.indptr = _iptr(self._k(_i32(I.indptr)))
        s.indices = _iptr(self._k(_i32(I.indices)))
        s.weights = _fptr(self._k(_f32(I.weights)))
        s.edge_delays = _iptr(self._k(_i32(I.edge_delays)))
        s.nnz = I.nnz
        s.delay_mode, s.single_delay = I.delay_mode, int(I.single_delay)
        s.max_delay_hint = int(I.max_delay_hint)
        s.bold_period = int(I.bold_period) if I.record_bold else 0
        s.bold_tr_ms = float(I.bold_tr_ms)
        s.step0 = int(I.step0)
        for name in ("history", "bold_state"):   # in/out buffers: written in place
            a = getattr(I, name)
            if a is not None:
                if not (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous):
                    raise ValueError(f"Instance.{name} must be a C-contiguous float32 array (written in place)")
                setattr(s, name, _fptr(self._k(a)))
        out = Output(tavg=np.zeros((I.n_nodes, nv), np.float32),
                     final_state=np.zeros((I.n_nodes, nv), np.float32))
        s.tavg = _fptr(out.tavg)
        s.final_state = _fptr(out.final_state)
        if I.record_traj:
            out.traj = np.zeros((I.n_steps, I.n_nodes, nv), np.float32)
            s.traj = _fptr(out.traj)
        if I.record_bold and I.bold_period > 0:
            rows = I.bold_rows()
            out.bold = np.zeros((max(rows, 1), I.n_nodes), np.float32)
            s.bold = _fptr(out.bold)
            s.bold_cap = rows
            s.bold_n = C.cast(self.bold_n.ctypes.data + i * 4, L.ip)
        self.outputs.append(out)

    def finish(self) -> list[Output]:
        for i, o in enumerate(self.outputs):
            if o.bold is not None:
                o.bold = o.bold[: int(self.bold_n[i])]
        return self.outputs


def _opts(device: int, exact: bool, graph: bool, block_size: int, grid: int, autotune: bool = False,
          keep_noise: bool = False) -> L.HbRunOpts:
    o = L.HbRunOpts()
    o.device = device
    o.flags = ((L.FLAG_EXACT if exact else 0) | (0 if graph else L.FLAG_NO_GRAPH) | (L.FLAG_AUTOTUNE if autotune else 0)
               | (L.FLAG_KEEP_NOISE if keep_noise else 0))
    o.block_size = block_size
    o.grid_waves = grid
    return o


class Plan:
    """An uploaded ragged batch: run() replays the captured graph; outputs stay
    in HBM until download()."""

    def __init__(self, insts: Sequence[Instance], device: int = 0, exact: bool = False,
                 graph: bool = True, block_size: int = 0, grid: int = 0, autotune: bool = False,
                 keep_noise: bool = False):
        """keep_noise: the first run() generates the noise, later runs reuse it
        (pre-generated noise resident in HBM, as an input)."""
        self._m = _Marshal(insts)
        self._h = C.c_void_p()
        o = _opts(device, exact, graph, block_size, grid, autotune, keep_noise)
        L.check(L.lib().hb_plan_create(self._m.arr, len(insts), C.byref(o), C.byref(self._h)))
        self.node_steps = float(sum(I.n_nodes * I.n_steps for I in insts))

    def run(self) -> L.HbRunStats:
        st = L.HbRunStats()
        L.check(L.lib().hb_plan_run(self._h, C.byref(st)))
        return st

    def download(self) -> list[Output]:
        L.check(L.lib().hb_plan_download(self._h))
        return self._m.finish()

    def close(self):
        if self._h:
            L.lib().hb_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def run_instances(insts: Sequence[Instance], device: int = 0, exact: bool = False,
                  graph: bool = True, block_size: int = 0, grid: int = 0, autotune: bool = False):
    """Run a ragged batch once; returns (outputs, stats)."""
    m = _Marshal(insts)
    st = L.HbRunStats()
    o = _opts(device, exact, graph, block_size, grid, autotune)
    L.check(L.lib().hb_run_instances(m.arr, len(insts), C.byref(o), C.byref(st)))
    return m.finish(), st


def autotune_entries() -> list:
    """Cached launch-shape choices: dicts of the shape class and its choice."""
    n = L.lib().hb_autotune_entries(None, 0)
    buf = np.zeros((max(n, 1), 8), np.int32)
    L.lib().hb_autotune_entries(buf.ctypes.data_as(L.ip), n)
    keys = ("model", "np", "n_nodes", "blocked", "lring", "integ_class", "stoch", "choice")
    return [dict(zip(keys, map(int, r))) for r in buf[:n]]


def philox_normal(seed: int, stream: int, n: int, device: int = 0) -> np.ndarray:
    """Device-generated Philox4x32-10 normals (the noise the planner uses)."""
    out = np.zeros(n, np.float32)
    L.check(L.lib().hb_philox_normal(seed, stream, n, _fptr(out), device))
    return out
