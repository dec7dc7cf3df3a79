"""`python -m hyburn_amd run`: the `hyburn run` command on the HIP backend.

Mirrors the reference command line (src/cli.rs:16-60, 785-842, 894-974, 1040-1103):

  run -c sim.toml -o output [-b hip] [--checkpoint FILE] [--resume FILE]
      [--sweep sweep.toml] [--noise-seed S]

* without --sweep: one plain-engine run (HybridEngine semantics), writing
  state_final_sub0.npy [nvar, nnodes, nmodes] and state_traj.npy
  [n_steps, nvar, nnodes, nmodes], var-major as the reference stores them;
* with --sweep: a batched sweep (BatchHybridEngine semantics) over the sweep
  file's parameter values, writing tavg_sub{i}.npy with the reference's header
  [n_sweep, nvar, nnodes] over its node-major [n_sweep, nnodes, nvar] data.

--resume restores a HYBURNCK checkpoint before the run and --checkpoint
saves one after it (runtime.rs:519-742, cli.rs:1057-1071): the reference's
version 2 file for a single network without a delay history to carry, the
batch files (versions 3 and 4) otherwise. Sweep mode ignores both, with a
warning, as the reference does (cli.rs:885-890).

n_steps = (sim_length / dt) truncated in f64, as the reference computes it.
Only the hip backend is provided: any other --backend value is an error
(there is no CPU path in this library).
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

from . import _lib as L
from .engine import (BatchHybridEngine, SimConfig, SweepConfig, SweepParam, _needs_multi, parse_sweep_param,
                     write_npy_f32)


def _backend(name: str) -> None:
    got = L.lib().hb_select_backend(name.encode())
    if got is None:
        raise SystemExit(f"backend '{name}' is not provided by this library (hip only): "
                         + L.lib().hb_last_error().decode(errors="replace"))


def run(config: str, output: str, backend: str = "hip", sweep: str | None = None, noise_seed: int = 0,
        device: int = 0, checkpoint: str | None = None, resume: str | None = None) -> dict:
    """Run the command; returns the written file paths."""
    _backend(backend)
    cfg = SimConfig.from_file(config)
    cfg.validate()
    n_steps = cfg.n_steps()
    os.makedirs(output, exist_ok=True)
    written = {}
    if sweep:
        if checkpoint:
            print("warning: Checkpoint ignored in sweep mode", file=sys.stderr)
        if resume:
            print("warning: Resume ignored in sweep mode", file=sys.stderr)
        sc = SweepConfig.from_file(sweep)
        values = sc.generate_values()
        if not values:
            raise SystemExit("Sweep config has no values")
        sub_idx, param_idx = parse_sweep_param(sc.parameter_name, cfg)
        eng = BatchHybridEngine.from_config(cfg, len(values), device, noise_seed=noise_seed)
        res = eng.run_sweep(SweepParam(sub_idx, param_idx), values, n_steps)
        eng.close()
        for i, t in enumerate(res.tavg):   # one file per subnetwork (cli.rs:958-970)
            sub = cfg.subnetworks[i]
            nnodes, nvar = sub.nnodes, L.MODEL_INFO[L.MODEL_NAMES[sub.model]][0]
            path = os.path.join(output, f"tavg_sub{i}.npy")
            # cli.rs:958-970 labels the [n_sweep, nnodes, nvar] data as [n_sweep, nvar, nnodes]
            write_npy_f32(path, t, [res.n_sweep, nvar, nnodes])
            written[f"tavg_sub{i}"] = path
    elif _needs_multi(cfg) or cfg.stimuli:
        # several subnetworks, links between them, cvar_map remaps, stimuli or
        # a model only the multi-subnetwork kernel integrates
        written.update(_plain_multi(cfg, n_steps, output, noise_seed, device, checkpoint, resume))
    else:
        # HybridEngine reads an inline initial state var-major [nvar][nnodes]
        # (construction.rs:557-565); the batch engine node-major (engine.rs:161-170)
        s0 = cfg.subnetworks[0]
        if isinstance(s0.initial_state, list):
            nv0 = L.MODEL_INFO[L.MODEL_NAMES[s0.model]][0]
            s0.initial_state = np.asarray(s0.initial_state, np.float32).reshape(nv0, s0.nnodes).T.ravel().tolist()
        eng = BatchHybridEngine.from_config(cfg, 1, device, engine=L.ENGINE_PLAIN, noise_seed=noise_seed)
        if resume:
            eng.resume(resume)
        # a plain run: no parameter is swept (an index the dfun ignores)
        res = eng.run_sweep_with_trajectory(SweepParam(0, -1), [0.0], n_steps)
        if checkpoint:
            eng.checkpoint(checkpoint)
            written["checkpoint"] = checkpoint
        eng.close()
        nnodes, nvar = eng.nnodes, eng.nvar
        fin = np.asarray(res.final_states[0], np.float32).reshape(nnodes, nvar)
        traj = np.asarray(res.trajectories[0], np.float32).reshape(n_steps, nnodes, nvar)
        p_fin = os.path.join(output, "state_final_sub0.npy")
        p_traj = os.path.join(output, "state_traj.npy")
        # HybridEngine state layout is [nvar, nnodes, nmodes] (runtime.rs, cli.rs:1073-1095)
        write_npy_f32(p_fin, fin.T.copy(), [nvar, nnodes, 1])
        write_npy_f32(p_traj, traj.transpose(0, 2, 1).copy(), [n_steps, nvar, nnodes, 1])
        written["state_final_sub0"] = p_fin
        written["state_traj"] = p_traj
    return written


def _plain_multi(cfg, n_steps: int, output: str, noise_seed: int, device: int, checkpoint: str | None = None,
                 resume: str | None = None) -> dict:
    """HybridEngine run of several subnetworks (cli.rs:1040-1103) through the
    plain mode of hb_run_multi: state_final_sub{i}.npy as [nvar, nnodes, 1]
    and state_traj.npy as the flat per-step concatenation of every
    subnetwork's [nvar, nnodes, nmodes] state (runtime.rs:430-434)."""
    from . import checkpoint as ck
    from .engine import _multi_from_config, multi_checkpoint_states
    from .multi import run_multi
    for s in cfg.subnetworks:   # inline initial states are var-major in HybridEngine
        if isinstance(s.initial_state, list):
            nv = L.MODEL_INFO[L.MODEL_NAMES[s.model]][0]
            s.initial_state = np.asarray(s.initial_state, np.float32).reshape(nv, s.nnodes).T.ravel().tolist()
    subs, links = _multi_from_config(cfg, plain=True)
    if cfg.integrator not in L.INTEGRATOR_NAMES:
        raise SystemExit(f"unknown integrator {cfg.integrator}")
    max_delay = max([max(p.delays) for p in cfg.projections if p.delays] or [0])
    stims = [(t["target"], t["temporal"].lower(), t["params"]) for t in cfg.stimuli]
    step0, state, hist, dt, integ, nsig = 0, None, None, cfg.dt, L.INTEGRATOR_NAMES[cfg.integrator], cfg.nsig
    if resume:   # runtime.rs:576-742: step, dt, integrator, nsig and states from the file
        c = ck.read(resume)
        state, hist = multi_checkpoint_states(c, subs, links, max_delay, 1)
        step0, dt, integ, nsig = int(c["step"]), float(c["dt"]), c["integrator"], c["nsig"]
    r = run_multi(subs, links, dt=dt, n_steps=n_steps, integrator=integ, plain=True, stimuli=stims, dt_ms=dt,
                  max_delay=max_delay, nsig=nsig, noise_seed=noise_seed, step0=step0, state=state, hist=hist,
                  record_trajectory=True, device=device)
    written = {}
    if checkpoint:
        ck.write_batch(checkpoint, step=step0 + n_steps, dt=dt, integrator=integ, nsig=nsig,
                       shapes=[(L.MODEL_INFO[S.model][0], S.n_nodes) for S in subs], states=r.final_states,
                       hist=r.hist)
        written["checkpoint"] = checkpoint
    for i, (S, f) in enumerate(zip(subs, r.final_states)):
        nv = L.MODEL_INFO[S.model][0]
        p = os.path.join(output, f"state_final_sub{i}.npy")
        write_npy_f32(p, f[0].T.copy(), [nv, S.n_nodes, 1])
        written[f"state_final_sub{i}"] = p
    traj = np.concatenate([np.concatenate([t[k, 0].T.ravel() for t in r.trajectories]) for k in range(n_steps)]) \
        if n_steps else np.zeros(0, np.float32)
    p = os.path.join(output, "state_traj.npy")
    write_npy_f32(p, traj, [traj.size])
    written["state_traj"] = p
    return written


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m hyburn_amd")
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run", help="run a simulation or a parameter sweep")
    r.add_argument("-c", "--config", default="sim.toml")
    r.add_argument("-o", "--output", default="output")
    r.add_argument("-b", "--backend", default="hip")
    r.add_argument("--checkpoint", default=None, help="save a HYBURNCK checkpoint after the run")
    r.add_argument("--resume", default=None, help="resume from a HYBURNCK checkpoint")
    r.add_argument("--sweep", default=None)
    r.add_argument("--noise-seed", type=int, default=0)
    r.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    try:
        out = run(a.config, a.output, a.backend, a.sweep, a.noise_seed, a.device, a.checkpoint, a.resume)
    except L.HyburnError as e:
        print(f"error: {e}", file=sys.stderr)
        return 1
    for k, v in out.items():
        print(f"{k}: {v}")
    return 0
