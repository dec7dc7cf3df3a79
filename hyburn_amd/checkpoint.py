"""HYBURNCK checkpoints (src/engine/runtime.rs:503-742), extended to batch engines.

Versions 1 and 2 are the reference's single-simulation files. The header is
magic "HYBURNCK", u64 version, u64 step, f64 dt, an integrator byte plus 7
padding bytes, then the nsig block. Version 1 holds one f32 plus 4 padding
bytes; version 2 holds u32 nsig_len, 4 padding bytes and nsig_len f32. Next
come u64 n_subnetworks, then (nvar, nnodes, nmodes) as u64 per subnetwork,
then each subnetwork's f32 state as [nvar][nnodes][nmodes]. Everything is
little-endian.

Version 3 is the batch extension written here. It has the same header and
shapes, then u64 n_sweep and u64 horizon H (0 without delays). Then come each
subnetwork's states [n_sweep][nnodes][nvar] and, when H > 0, the coupling
variable history [n_sweep][total nodes][H][2] indexed by step modulo H. That
is everything a multi-subnetwork batch run needs to continue bit for bit.

Version 4 is the single-network engine's file. It has the same header and the
one shape, then u64 n_sweep and u64 hist_floats (per sweep point, 0 without
delays), the states [n_sweep][nnodes][nvar] and, when hist_floats > 0, each
point's delay history as the engine keeps it: [max_delay][nnodes][channels],
oldest step first (hb_history_floats). A single-network engine writes
version 2, the reference's own format, whenever that loses nothing: one
sweep point and no history to carry.
"""
from __future__ import annotations

import struct

import numpy as np

from . import _lib as L

MAGIC = b"HYBURNCK"
# runtime.rs:531-538
INTEGRATOR_BYTE = {L.HEUN: 1, L.EULER: 2, L.EULER_STOCH: 3, L.HEUN_STOCH: 4, L.RK4: 5, L.RK4_STOCH: 6}
BYTE_INTEGRATOR = {v: k for k, v in INTEGRATOR_BYTE.items()}


def _bad(msg: str):
    raise L.HyburnError(L.HB_INVALID_STATE, msg)


def _header(f, version: int, step: int, dt: float, integrator: int, nsig) -> None:
    nsig = np.atleast_1d(np.asarray(nsig, np.float32))
    f.write(MAGIC)
    f.write(struct.pack("<QQd", version, int(step), float(dt)))
    f.write(bytes([INTEGRATOR_BYTE[integrator]]) + bytes(7))
    f.write(struct.pack("<I", len(nsig)) + bytes(4))
    f.write(nsig.astype("<f4").tobytes())


def write_batch(path: str, *, step: int, dt: float, integrator: int, nsig, shapes, states, hist=None) -> None:
    """Version 3. shapes: (nvar, nnodes) per subnetwork; states: [n_sweep][nnodes][nvar] per subnetwork."""
    n_sweep = int(np.asarray(states[0]).shape[0])
    H = 0 if hist is None else int(np.asarray(hist).shape[2])
    with open(path, "wb") as f:
        _header(f, 3, step, dt, integrator, nsig)
        f.write(struct.pack("<Q", len(shapes)))
        for nv, nn in shapes:
            f.write(struct.pack("<QQQ", nv, nn, 1))
        f.write(struct.pack("<QQ", n_sweep, H))
        for (nv, nn), st in zip(shapes, states):
            a = np.ascontiguousarray(st, "<f4")
            if a.shape != (n_sweep, nn, nv):
                raise ValueError("state shape does not match the subnetwork")
            f.write(a.tobytes())
        if H:
            f.write(np.ascontiguousarray(hist, "<f4").tobytes())


def write_reference(path: str, *, step: int, dt: float, integrator: int, nsig, nvar: int, nnodes: int,
                    state) -> None:
    """Version 2 exactly as HybridEngine::checkpoint writes it (runtime.rs:519-561)
    for one subnetwork: state [nnodes][nvar] node-major in, stored
    [nvar][nnodes][nmodes = 1]."""
    a = np.asarray(state, np.float32)
    if a.shape != (nnodes, nvar):
        raise ValueError("state shape does not match the subnetwork")
    with open(path, "wb") as f:
        _header(f, 2, step, dt, integrator, nsig)
        f.write(struct.pack("<QQQQ", 1, nvar, nnodes, 1))
        f.write(np.ascontiguousarray(a.T, "<f4").tobytes())


def write_single(path: str, *, step: int, dt: float, integrator: int, nsig, nvar: int, nnodes: int, states,
                 hist=None) -> None:
    """Version 4: states [n_sweep][nnodes][nvar], hist [n_sweep][hist_floats] or None."""
    a = np.ascontiguousarray(states, "<f4")
    n = a.shape[0]
    if a.shape != (n, nnodes, nvar):
        raise ValueError("state shape does not match the subnetwork")
    h = None if hist is None else np.ascontiguousarray(hist, "<f4")
    if h is not None and (h.ndim != 2 or h.shape[0] != n):
        raise ValueError("history must be [n_sweep][hist_floats]")
    with open(path, "wb") as f:
        _header(f, 4, step, dt, integrator, nsig)
        f.write(struct.pack("<QQQQ", 1, nvar, nnodes, 1))
        f.write(struct.pack("<QQ", n, 0 if h is None else h.shape[1]))
        f.write(a.tobytes())
        if h is not None:
            f.write(h.tobytes())


def read(path: str) -> dict:
    """Parse a version 1 to 4 file. Returns step, dt, integrator, nsig,
    shapes [(nvar, nnodes, nmodes)], n_sweep, states (per subnetwork
    [n_sweep][nnodes][nvar]; versions 1-2 as n_sweep = 1, converted from
    [nvar][nnodes][nmodes] with nmodes = 1), hist (version 3's
    multi-subnetwork history, or None) and hist_single (version 4's
    [n_sweep][hist_floats], or None)."""
    try:
        buf = open(path, "rb").read()
    except OSError as e:
        raise L.HyburnError(L.HB_IO_ERROR, str(e))
    off = 8

    def take(n):
        nonlocal off
        if off + n > len(buf):
            _bad("Checkpoint data truncated: expected more bytes")
        b = buf[off:off + n]
        off += n
        return b

    if len(buf) < 8 or buf[:8] != MAGIC:
        _bad("Invalid checkpoint file (bad magic)")
    version, step = struct.unpack("<QQ", take(16))
    (dt,) = struct.unpack("<d", take(8))
    ib = take(8)[0]
    if version == 1:
        (v,) = struct.unpack("<f", take(4))
        take(4)
        nsig = np.array([v], np.float32)
    elif version in (2, 3, 4):
        (n,) = struct.unpack("<I", take(4))
        take(4)
        nsig = np.frombuffer(take(4 * n), "<f4").astype(np.float32)
    else:
        _bad(f"Unsupported checkpoint version: {version} (expected 1, 2, 3 or 4)")
    if ib not in BYTE_INTEGRATOR:
        _bad(f"Unknown integrator kind in checkpoint: {ib}")
    (n_subs,) = struct.unpack("<Q", take(8))
    shapes = [struct.unpack("<QQQ", take(24)) for _ in range(n_subs)]
    out = {"version": version, "step": step, "dt": dt, "integrator": BYTE_INTEGRATOR[ib], "nsig": nsig,
           "shapes": shapes, "hist": None, "hist_single": None}
    if version == 4:
        if n_subs != 1:
            _bad("a version 4 checkpoint holds one subnetwork")
        n_sweep, hf = struct.unpack("<QQ", take(16))
        nv, nn, nm = shapes[0]
        out["n_sweep"] = n_sweep
        out["states"] = [np.frombuffer(take(4 * n_sweep * nn * nv), "<f4").astype(np.float32).reshape(n_sweep, nn, nv)]
        if hf:
            out["hist_single"] = np.frombuffer(take(4 * n_sweep * hf), "<f4").astype(np.float32).reshape(n_sweep, hf)
    elif version == 3:
        n_sweep, H = struct.unpack("<QQ", take(16))
        out["n_sweep"] = n_sweep
        out["states"] = [np.frombuffer(take(4 * n_sweep * nn * nv), "<f4").astype(np.float32).reshape(n_sweep, nn, nv)
                         for (nv, nn, nm) in shapes]
        if H:
            n_tot = sum(nn for (_, nn, _) in shapes)
            out["hist"] = np.frombuffer(take(4 * n_sweep * n_tot * H * 2), "<f4").astype(np.float32).reshape(
                n_sweep, n_tot, H, 2)
    else:
        out["n_sweep"] = 1
        sts = []
        for (nv, nn, nm) in shapes:
            a = np.frombuffer(take(4 * nv * nn * nm), "<f4").astype(np.float32).reshape(nv, nn, nm)
            if nm != 1:
                _bad("nmodes > 1 is not supported by this backend")
            sts.append(a[:, :, 0].T[None].copy())
        out["states"] = sts
    return out
