#!/usr/bin/env python3
"""Benchmark of the batched neural-mass integration path on one MI355X.

One step = one full pass of the hot path over the batched-sweep workload:
Generic2dOscillator, 74 nodes x 128 sweep points (Linear coupling G
log-uniform in [1e-3, 1e-1] x parameter a in U[-2, 2]), symmetric 20%
lognormal connectome with per-edge delays <= 500 steps, Heun-stochastic
dt = 0.1 ms, 10,000 integration steps, nsig = 0.01 Philox noise generated on
device before timing (pre-generated, as the workload specifies). Inputs are
resident in HBM before timing; the timed region covers the fused integration
kernel; the noise generation time is reported beside it.

Prints one JSON line (metric, value, unit, roofline, cpu_baseline, ...). Extra
legs, each with its own switch, add objects to the same line: "ragged" (4096
ragged MPR/Jansen-Rit instances in one plan), "features" (FC, FCD and spectral
features of 1024 BOLD-length series), "ring" (the G2DO -> JR -> WC network of
the reference's generic-engine benchmark), "dense" (dense all-to-all coupling
of 4096 points on the matrix cores against per-lane row sums) and "adjoint"
(gradients of the FC fit and Kuramoto order-parameter workloads).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic FP32 work per node-step of this workload (Heun: 2 G2DO dfuns of
# 20 FLOP, 2 x 7 FLOP of stochastic Heun update per state variable, 2 FLOP per
# CSR edge at 14.6 edges/row, 2 FLOP Linear post): 85 FLOP.
FLOP_PER_NODE_STEP = 2 * 20 + 2 * 7 + 2 * 14.6 + 2
# Algorithmic HBM bytes per node-step of the forward kernel: the f32 noise of
# 2 state variables, read once (the delay ring lives in LDS; tavg/final
# outputs, 2 x 4 B per node and variable, are added per launch).
BYTES_PER_NODE_STEP = 2 * 4
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def read_traffic(name: str = "forward_sweep_g2do_74x128_traffic.txt"):
    """HBM bytes per launch of a kernel from a committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    vals = {}
    for line in open(p):
        if ":" in line and not line.lstrip().startswith("#"):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def host_cpu() -> dict:
    """The host the CPU baseline runs on: logical CPUs, the ones this process
    may use, and the CPU model string."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "usable": usable, "cpu_model": model}


def cpu_threads(args) -> int:
    """Threads of the CPU baseline: every CPU this process may use, up to 16,
    the host-CPU share of one GPU on these machines (worker pools are sized to
    it whatever nproc shows)."""
    return args.cpu_threads or min(host_cpu()["usable"], 16)


def ragged_bytes(insts):
    """Algorithmic HBM bytes of the ragged forward: per node-step, a ring kept in
    HBM costs one 4-byte write per channel plus one 4-byte read per CSR edge and
    channel, 4*C*(1 + nnz/N) with no cache credit ("no-reuse"); the compulsory
    part is 4*C*2 (each value written and read once). Rings of this workload
    (68-400 nodes x up to 801 slots) exceed LDS, so every instance's ring is in
    HBM. Outputs (tavg, final state, BOLD rows) are added per instance."""
    no_reuse = compulsory = 0.0
    for I in insts:
        C = 1   # MPR with cv = 0 and Jansen-Rit store one ring channel
        ns = I.n_nodes * I.n_steps
        no_reuse += ns * 4 * C * (1 + I.nnz / I.n_nodes)
        compulsory += ns * 4 * C * 2
        out = 2 * 4 * I.n_nodes * I.nvar + 4 * I.n_nodes * I.bold_rows()
        no_reuse += out
        compulsory += out
    return no_reuse, compulsory


def cpu_baseline(insts, threads: int):
    """The oracle (a C restatement of the reference NdArray path) on host cores."""
    from oracle import hb_oracle
    secs = hb_oracle.run_many_f32(insts, n_threads=threads)
    ns = sum(I.n_nodes * I.n_steps for I in insts)
    return ns / secs, secs


def copy_calibration(mib: int) -> None:
    """One device-to-device copy of `mib` MiB: the reference dispatch for
    scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes (read mib, write mib)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0        # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[bench] copy calibration: {mib} MiB device-to-device")


def ragged_leg(args):
    """Ragged 4096-instance MPR/JR RK4 sweep: one persistent launch vs
    per-instance plain launches of the same kernel (sampled)."""
    from hyburn_amd.instances import Plan
    from hyburn_amd.synth import ragged_mpr_jr
    insts = ragged_mpr_jr(n_inst=args.ragged_n, t_scale=args.ragged_tscale)
    ns = sum(I.n_nodes * I.n_steps for I in insts)
    log(f"[bench] ragged: {len(insts)} instances, {ns:.3e} node-steps; planning")
    from hyburn_amd.instances import autotune_entries
    from hyburn_amd import _lib as L
    t0 = time.perf_counter()
    reps = max(1, args.ragged_steps)
    with Plan(insts, autotune=True) as P:
        plan_s = time.perf_counter() - t0
        P.run()
        L.check(L.lib().hb_device_synchronize(0))
        k_ms = []
        t0 = time.perf_counter()
        for _ in range(reps):
            k_ms.append(P.run().kernel_ms)
        L.check(L.lib().hb_device_synchronize(0))
        wall = (time.perf_counter() - t0) / reps
        outs = P.download()
    assert all(np.all(np.isfinite(o.final_state)) for o in outs), "non-finite ragged output"
    k = sum(k_ms) / len(k_ms)
    no_reuse, compulsory = ragged_bytes(insts)
    achieved = no_reuse / (k * 1e-3) / 1e9
    res = {"n_instances": len(insts), "node_steps": ns, "value": ns / wall, "wall_s": wall, "repeats": reps,
           "kernel_ms": k, "plan_and_autotune_s": plan_s, "autotuned": autotune_entries(),
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": achieved / PEAK_HBM_GBS, "traffic": read_traffic("forward_ragged_4096_traffic.txt"),
                        "algorithmic_bytes": no_reuse, "compulsory_bytes": compulsory,
                        "compulsory_frac": compulsory / (k * 1e-3) / 1e9 / PEAK_HBM_GBS,
                        "kernel": "hb::fast::k_single<MPR|JR,1,hbm_ring> launch groups (one hipGraph)",
                        "note": "algorithmic bytes = ring write + one read per CSR edge and channel, no cache "
                                "credit, per node-step (4*C*(1+nnz/N)); achieved over the graph's HIP-event time"}}
    # per-instance plain launches: one plan (one launch) per instance, sampled
    k = min(args.ragged_plain_sample, len(insts))
    idx = list(range(0, len(insts), max(1, len(insts) // k)))[:k] if k > 0 else []
    sub_ns, sub_t = 0.0, 0.0
    for i in idx:
        with Plan([insts[i]]) as P:
            P.run()
            t0 = time.perf_counter()
            P.run()
            sub_t += time.perf_counter() - t0
            sub_ns += insts[i].n_nodes * insts[i].n_steps
    if idx:
        res["plain_launch_value"] = sub_ns / sub_t
        res["plain_launch_sample"] = len(idx)
        res["speedup_vs_plain_launches"] = res["value"] / res["plain_launch_value"]
    if not args.no_cpu:
        # bounded CPU sample: every 64th instance at full length, all host threads
        threads = cpu_threads(args)
        sample = insts[:: max(1, len(insts) // 64)]
        v, secs = cpu_baseline(sample, threads)
        res["cpu_baseline"] = {"value": v, "unit": "node-steps/s", "cores": threads, "kind": "port", **host_cpu(),
                               "sample": f"{len(sample)} of the {len(insts)} instances (every "
                                         f"{max(1, len(insts) // 64)}th, full length, BOLD monitor on as on "
                                         f"the GPU), {secs:.1f} s wall"}
        res["speedup_vs_cpu"] = res["value"] / v
    return res


def features_leg(args):
    """Batched FC, FCD (fc_matrix / fcd_matrix) and spectral features
    (spectral_features of every node series) of many BOLD-length series:
    1024 series x 1200 samples x 84 nodes; FCD windows of 60 samples every 6;
    Welch PSD with 2048-point transforms of 256-sample windows at fs = 1 kHz.
    The C ABI takes host buffers, so the times include the host-device copies."""
    from hyburn_amd.features import fc_batch, fcd_batch, spectral_batch
    rng = np.random.default_rng(0)
    B, T, N, Lw, S = args.features_series, 1200, 84, 60, 6
    X = (np.cumsum(rng.normal(size=(B, T, N)), axis=1) + np.cumsum(rng.normal(size=(B, T, 1)), axis=1)).astype(
        np.float32)
    res = {"workload": f"{B} series x {T} samples x {N} nodes, FCD window {Lw} stride {S}"}
    for name, fn in (("fc", lambda a: fc_batch(a)), ("fcd", lambda a: fcd_batch(a, Lw, S)),
                     ("spectral", lambda a: spectral_batch(a, 1000.0))):
        fn(X[:8])
        t0 = time.perf_counter()
        fn(X)
        wall = time.perf_counter() - t0
        res[name] = {"series_per_s": B / wall, "call_ms": wall * 1e3}
    if not args.no_cpu:
        from oracle import hb_oracle
        k = 4
        t0 = time.perf_counter()
        for b in range(k):
            hb_oracle.fc(X[b])
        fc_s = (time.perf_counter() - t0) / k
        t0 = time.perf_counter()
        for b in range(k):
            hb_oracle.fcd(X[b], Lw, S)
        fcd_s = (time.perf_counter() - t0) / k
        t0 = time.perf_counter()
        for n in range(N):
            hb_oracle.spectral_features(X[0, :, n], 1000.0)
        sp_s = time.perf_counter() - t0
        res["cpu_baseline"] = {"fc_series_per_s": 1.0 / fc_s, "fcd_series_per_s": 1.0 / fcd_s,
                               "spectral_series_per_s": 1.0 / sp_s, "cores": 1, "kind": "port",
                               "sample": f"{k} of the series (FC, FCD) and the {N} nodes of one series "
                                         f"(spectral, numpy FFT), one host thread"}
        res["fc"]["speedup_vs_cpu"] = res["fc"]["series_per_s"] * fc_s
        res["fcd"]["speedup_vs_cpu"] = res["fcd"]["series_per_s"] * fcd_s
        res["spectral"]["speedup_vs_cpu"] = res["spectral"]["series_per_s"] * sp_s
    return res


def ring_leg(args):
    """The reference's generic-engine benchmark network: G2DO -> JR -> WC ring,
    Scalar all-to-all Linear coupling w = 0.01, zero delays, Heun with
    hybrid_integrator (Heun for G2DO, Euler for JR and WC), zero initial
    state, G2DO I swept over linspace(-0.5, 0.5, n_sweep), 1000 steps of
    dt = 0.1. Published: 360 ms (76 nodes) and 341 ms (164 nodes) per 1024
    points on an RTX 4090 (6.49e8 and 1.48e9 node-steps/s)."""
    from hyburn_amd import _lib as L
    from hyburn_amd.multi import Link, Subnet, run_multi
    from hyburn_amd.synth import G2DO_DEFAULT, JR_DEFAULT, WC_DEFAULT
    published = {76: 6.49e8, 164: 1.48e9}
    res = {}
    for n in (76, 164):
        subs = [Subnet(L.G2DO, n, G2DO_DEFAULT), Subnet(L.JANSEN_RIT, n, JR_DEFAULT),
                Subnet(L.WILSON_COWAN, n, WC_DEFAULT)]
        links = [Link(0, 1, 0.01), Link(1, 2, 0.01), Link(2, 0, 0.01)]
        B, T = 1024, 1000
        vals = (-0.5 + np.arange(B, dtype=np.float32) / np.float32(B - 1)).astype(np.float32)
        kw = dict(dt=0.1, n_steps=T, integrator=L.HEUN, hybrid_integrator=True, n_sweep=B, sweep=(0, 1, vals))
        run_multi(subs, links, **kw)
        t0 = time.perf_counter()
        r = run_multi(subs, links, **kw)
        wall = time.perf_counter() - t0
        ns = B * 3 * n * T
        res[f"ring_{n}"] = {"node_steps_per_s": ns / (r.kernel_ms * 1e-3), "kernel_ms": r.kernel_ms,
                            "call_ms": wall * 1e3, "published_rtx4090": published[n],
                            "vs_published": ns / (r.kernel_ms * 1e-3) / published[n]}
    res["workload"] = "G2DO->JR->WC ring, 1024 points x 1000 steps, hybrid Heun/Euler, Scalar w=0.01"
    return res


def dense_leg(args):
    """Dense all-to-all coupling at large batch: Generic2dOscillator, 74 nodes,
    the batched sweep's 20% lognormal connectome as a dense matrix, no delay,
    Heun, 4096 sweep points of I x 2000 steps. The matrix-core product over
    16-instance tiles against the per-lane row sums (HB_FLAG_NO_MFMA), both
    FMA mode, inputs resident, kernel time from HIP events."""
    import ctypes as C

    from hyburn_amd import _lib as L
    from hyburn_amd.instances import Instance, _Marshal, _opts
    from hyburn_amd.synth import G2DO_DEFAULT, lognormal_connectome
    N, B, T = 74, args.dense_batch, 2000
    W = lognormal_connectome(np.random.default_rng(3), N, 0.20)
    x0 = np.random.default_rng(5).normal(0.0, 0.1, (N, 2)).astype(np.float32)
    insts = [Instance(model=L.G2DO, n_nodes=N, n_steps=T, dt=0.1, params=G2DO_DEFAULT.copy(), x0=x0,
                      integrator=L.HEUN, wkind=L.W_DENSE, dense=W, cfn=L.CF_LINEAR, cparams=(0.05, 0.0),
                      sweep_form=1, sweep_val=float(v)) for v in np.linspace(-0.5, 0.5, B)]
    res = {"workload": f"G2DO {N} nodes, dense 20% lognormal, no delay, Heun, {B} points x {T} steps"}
    ns = float(B * N * T)
    for name, extra in (("matrix_cores", 0), ("lane_row_sums", L.FLAG_NO_MFMA)):
        m = _Marshal(insts)
        o = _opts(0, False, True, 0, 0)
        o.flags |= extra
        h = C.c_void_p()
        L.check(L.lib().hb_plan_create(m.arr, len(insts), C.byref(o), C.byref(h)))
        st = L.HbRunStats()
        L.check(L.lib().hb_plan_run(h, C.byref(st)))
        ks = []
        for _ in range(3):
            L.check(L.lib().hb_plan_run(h, C.byref(st)))
            ks.append(st.kernel_ms)
        L.lib().hb_plan_destroy(h)
        k = sum(ks) / len(ks)
        res[name] = {"node_steps_per_s": ns / (k * 1e-3), "kernel_ms": k, "block": st.block_size}
    res["speedup"] = res["matrix_cores"]["node_steps_per_s"] / res["lane_row_sums"]["node_steps_per_s"]
    return res


def _fc64(B):
    n = B.shape[0]
    A = B - B.mean(0)
    sd = np.sqrt((A * A).sum(0) / n)
    sd = np.where(sd < 1e-12, 1.0, sd)
    return np.clip((A.T @ A) / n / (sd[:, None] * sd[None, :]), -1.0, 1.0)


_PLAIN: dict = {}


def adjoint_leg(name, insts, losses, checkpoint_every, n_params):
    """Gradient evaluations (forward with checkpoints + loss + reverse sweep)
    per second over the whole batch, against the plain forward of the same
    batch; central differences would need 2 * n_params plain forwards."""
    from hyburn_amd.adjoint import AdjointPlan
    from hyburn_amd.instances import Plan
    B = len(insts)
    log(f"[bench] adjoint {name}: {B} x {insts[0].n_nodes} nodes x {insts[0].n_steps} steps; planning")
    t0 = time.perf_counter()
    P = AdjointPlan(insts, losses, checkpoint_every=checkpoint_every)
    plan_s = time.perf_counter() - t0
    P.forward()
    P.backward()
    t0 = time.perf_counter()
    P.forward()
    P.backward()
    wall = time.perf_counter() - t0
    fwd_ms, bwd_ms, dev_bytes = P.stats.forward_ms, P.stats.backward_ms, P.stats.device_bytes
    P.close()
    key = name   # the workload: the plain forward of a batch is timed once for all its modes
    if key not in _PLAIN:
        with Plan(insts) as Q:
            Q.run()
            t0 = time.perf_counter()
            st = Q.run()
            _PLAIN[key] = (st.total_ms, time.perf_counter() - t0)
    plain_ms, plain_wall = _PLAIN[key]
    return {
        "workload": name, "n_instances": B, "n_nodes": insts[0].n_nodes, "n_steps": insts[0].n_steps,
        "grads_per_s": B / wall, "wall_ms": wall * 1e3, "adjoint_forward_ms": fwd_ms, "adjoint_backward_ms": bwd_ms,
        "plain_forward_ms": plain_ms, "cost_vs_forward": (fwd_ms + bwd_ms) / plain_ms,
        "checkpoint_every": checkpoint_every or "auto: whole trajectory kept when it fits a quarter of free HBM",
        "device_bytes": dev_bytes, "plan_s": plan_s,
        "fd_grads_per_s_est": B / (2 * n_params * plain_wall),
        "fd_note": f"central differences over {n_params} parameters per instance = {2 * n_params} plain forwards "
                   "per gradient, estimated from the measured plain forward",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-sweep", type=int, default=128)
    ap.add_argument("--n-steps", type=int, default=10_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-ragged", dest="ragged", action="store_false",
                    help="skip the 4096-instance ragged sweep leg")
    ap.add_argument("--ragged-n", type=int, default=4096)
    ap.add_argument("--ragged-tscale", type=float, default=1.0)
    ap.add_argument("--ragged-steps", type=int, default=3, help="timed passes of the ragged plan")
    ap.add_argument("--ragged-plain-sample", type=int, default=16)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the gradient legs")
    ap.add_argument("--no-ring", dest="ring", action="store_false",
                    help="skip the three-subnetwork ring leg")
    ap.add_argument("--no-features", dest="features", action="store_false",
                    help="skip the FC / FCD / spectral leg")
    ap.add_argument("--features-series", type=int, default=1024)
    ap.add_argument("--no-dense", dest="dense", action="store_false",
                    help="skip the dense all-to-all (matrix cores vs row sums) leg")
    ap.add_argument("--dense-batch", type=int, default=4096)
    ap.add_argument("--adj4-batch", type=int, default=256)
    ap.add_argument("--adj5-batch", type=int, default=512)
    ap.add_argument("--adj5-steps", type=int, default=10_000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run one device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path "
              "(instances would shard across GPUs without communication)", file=sys.stderr)
        return 2

    from hyburn_amd import _lib as L
    from hyburn_amd.instances import Plan
    from hyburn_amd.synth import sweep_g2do_74

    lib = L.lib()
    if lib.hb_device_count() < 1:
        print("no HIP device", file=sys.stderr)
        return 1
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    insts = sweep_g2do_74(n_sweep=args.n_sweep, n_steps=args.n_steps)
    node_steps = sum(I.n_nodes * I.n_steps for I in insts)
    log(f"[bench] workload: {len(insts)} x {insts[0].n_nodes} nodes x {insts[0].n_steps} steps")
    # the workload's noise is pre-generated (Philox, seeded): an input resident
    # in HBM like the connectome, generated by the first (untimed) run
    P = Plan(insts, keep_noise=True)
    first = P.run()
    noise_ms = first.total_ms - first.kernel_ms
    for _ in range(args.warmup):
        P.run()
    L.check(lib.hb_device_synchronize(0))
    kernel_ms, total_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = P.run()
        kernel_ms.append(st.kernel_ms)
        total_ms.append(st.total_ms)
    L.check(lib.hb_device_synchronize(0))
    wall = time.perf_counter() - t0
    outs = P.download()
    P.close()
    assert all(np.all(np.isfinite(o.tavg)) for o in outs), "non-finite output"

    ms_per_step = wall / args.steps * 1e3
    value = node_steps * args.steps / wall
    k_ms = sum(kernel_ms) / len(kernel_ms)
    flops = FLOP_PER_NODE_STEP * node_steps / (k_ms * 1e-3) / 1e12
    # The roofline that binds: bytes/BW (8 B / 8 TB/s) exceeds FLOP/peak
    # (85 FLOP / 157.3 TFLOP/s) per node-step, so the kernel is priced in HBM bytes.
    alg_bytes = BYTES_PER_NODE_STEP * node_steps + 2 * 4 * sum(I.n_nodes * I.nvar for I in insts)
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9
    traffic = read_traffic()
    line = {
        "metric": "node-steps/s (ragged sweep, B=128..4096) + adjoint grads/s vs plain & FD",
        "value": value,
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded connectome, delays, parameters; Philox noise)",
        "config": {
            "workload": "batched sweep: Generic2dOscillator 74 nodes x 128 points x 10k Heun-stochastic "
                        "steps, 20% lognormal CSR with per-edge delays <= 500 steps, nsig 0.01",
            "n_sweep": len(insts), "n_nodes": insts[0].n_nodes, "n_steps": insts[0].n_steps,
            "parallelism": "1 workgroup per sweep point, delay ring in LDS",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": traffic,
            "algorithmic_bytes": alg_bytes,
            "kernel": "hb::fast::k_single<G2DO,1,lds_ring>",
            "kernel_ms": k_ms,
            "fp32_tflops": flops,
            "fp32_frac": flops / PEAK_FP32_TFLOPS,
            "flop_per_node_step": FLOP_PER_NODE_STEP,
            "note": "each sweep point is a chain of 10k dependent steps: the kernel is latency bound, "
                    "far below both the HBM and the FP32 roofline",
        },
        "timing": {"kernel_ms": k_ms, "pass_ms": sum(total_ms) / len(total_ms),
                   "noise_generation_ms": noise_ms,
                   "with_noise_generation_node_steps_per_s": node_steps / ((ms_per_step + noise_ms) * 1e-3),
                   "note": "the noise (758 MB of Philox normals) is pre-generated, as the workload specifies, "
                           "before the timed region; generating it on the device takes noise_generation_ms"},
    }
    if not args.no_cpu:
        # bounded sample: the whole workload on all host threads (~1-3 s wall)
        threads = cpu_threads(args)
        from oracle import hb_oracle
        hb_oracle.lib()
        sample = sweep_g2do_74(n_sweep=args.n_sweep, n_steps=args.n_steps)
        v, secs = cpu_baseline(sample, threads)
        line["cpu_baseline"] = {"value": v, "unit": "node-steps/s", "cores": threads, "kind": "port", **host_cpu(),
                                "sample": f"the full workload ({len(sample)} points x {args.n_steps} steps) "
                                          f"on {threads} host threads, {secs:.2f} s wall "
                                          f"({secs * threads:.1f} core-s)"}
    if args.ragged:
        line["ragged"] = ragged_leg(args)
    if args.features:
        line["features"] = features_leg(args)
    if args.ring:
        line["ring"] = ring_leg(args)
    if args.dense:
        line["dense"] = dense_leg(args)
    if not args.no_adjoint:
        import dataclasses
        from hyburn_amd.adjoint import Loss
        from hyburn_amd.instances import run_instances
        from hyburn_amd.synth import fc_fit_mpr_84, kuramoto_order_1000
        fc_insts, hidden = fc_fit_mpr_84(n_inst=args.adj4_batch)
        (h,), _ = run_instances([dataclasses.replace(hidden, record_bold=True)])
        fc_t = _fc64(h.bold.astype(np.float64)).astype(np.float32)
        fc_name = "MPR 84 nodes, 20k Heun steps, FC of native BOLD vs hidden-parameter FC; grads of G and 84 eta"
        fc_losses = [Loss(L.LOSS_FC, fc_t, 2) for _ in fc_insts]
        # the workload's own mode: a checkpoint every 500 steps, segments
        # recomputed in the reverse sweep; then the whole trajectory kept
        line["adjoint"] = {
            "fc_mpr_84": adjoint_leg(fc_name, fc_insts, fc_losses, 500, 1 + fc_insts[0].n_nodes),
            "fc_mpr_84_trajectory_kept": adjoint_leg(fc_name, fc_insts, fc_losses, 0, 1 + fc_insts[0].n_nodes),
        }
        del fc_insts
        kur_insts, tgt = kuramoto_order_1000(n_inst=args.adj5_batch, n_steps=args.adj5_steps)
        kur_name = "Kuramoto 1000 nodes, 5% CSR, delays <= 40 ms, RK4, order-parameter MSE; grads of G and 1000 omega"
        kur_losses = [Loss(L.LOSS_ORDER, tgt, 0) for _ in kur_insts]
        sqrt_t = int(np.ceil(np.sqrt(kur_insts[0].n_steps)))
        line["adjoint"]["order_kuramoto_1000"] = adjoint_leg(kur_name, kur_insts, kur_losses, sqrt_t,
                                                             1 + kur_insts[0].n_nodes)
        line["adjoint"]["order_kuramoto_1000_trajectory_kept"] = adjoint_leg(kur_name, kur_insts, kur_losses, 0,
                                                                             1 + kur_insts[0].n_nodes)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
