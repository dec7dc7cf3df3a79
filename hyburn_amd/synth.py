"""Seeded synthetic workloads (connectomes, delays, parameters, noise).

Noise: Philox4x32-10 (Salmon et al., SC'11) keyed by the 64-bit seed with
counter (pair index, stream id); two 53-bit uniforms per block, Box-Muller in
f64, cast to f32. `philox_normal` is the numpy statement of the stream the HIP
noise kernel (csrc/hb_noise.hip) generates.

Structure (connectomes, tract lengths, parameters, initial states) comes from
numpy's PCG64 `default_rng(seed)`; the same image gives the same arrays.
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib as L
from .instances import Instance

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def _philox4x32_10(c0, c1, c2, c3, k0: int, k1: int):
    for _ in range(10):
        p0 = c0.astype(np.uint64) * _M0
        p1 = c2.astype(np.uint64) * _M1
        hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & _MASK).astype(np.uint32)
        hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & _MASK).astype(np.uint32)
        c0, c1, c2, c3 = hi1 ^ c1 ^ np.uint32(k0), lo1, hi0 ^ c3 ^ np.uint32(k1), lo0
        k0 = (k0 + _W0) & 0xFFFFFFFF
        k1 = (k1 + _W1) & 0xFFFFFFFF
    return c0, c1, c2, c3


def philox_normal(seed: int, stream: int, n: int) -> np.ndarray:
    """Element k of stream `stream` under key `seed`, as f32 (see module doc)."""
    npairs = (n + 1) // 2
    p = np.arange(npairs, dtype=np.uint64)
    c0 = (p & _MASK).astype(np.uint32)
    c1 = (p >> np.uint64(32)).astype(np.uint32)
    c2 = np.full(npairs, stream & 0xFFFFFFFF, dtype=np.uint32)
    c3 = np.full(npairs, (stream >> 32) & 0xFFFFFFFF, dtype=np.uint32)
    c0, c1, c2, c3 = _philox4x32_10(c0, c1, c2, c3, seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    a = ((c0 >> np.uint32(5)).astype(np.uint64) << np.uint64(26)) | (c1 >> np.uint32(6)).astype(np.uint64)
    b = ((c2 >> np.uint32(5)).astype(np.uint64) << np.uint64(26)) | (c3 >> np.uint32(6)).astype(np.uint64)
    u1 = (a + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = b.astype(np.float64) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    th = 6.283185307179586 * u2
    z = np.empty(2 * npairs, dtype=np.float64)
    z[0::2] = r * np.cos(th)
    z[1::2] = r * np.sin(th)
    return z[:n].astype(np.float32)


# ---------------------------------------------------------------- structure --

G2DO_DEFAULT = np.array([1, 0, -2, -10, 0, 0.02, 3, 1, 0, 1, 1, 1], np.float32)     # g2do.rs:76-78
MPR_DEFAULT = np.array([1, 1, -5, 15, 0, 1, 0], np.float32)                        # mpr.rs:61-63
JR_DEFAULT = np.array([3.25, 22, 0.1, 0.05, 5.52, 0.0025, 0.56, 135, 1, 0.8, 0.25, 0.25, 0.22],
                      np.float32)                                                  # jansen_rit.rs:77-82
WC_DEFAULT = np.array([12, 4, 13, 11, 10, 10, 1.2, 2.8, 1, 0, 1, 4, 1, 0, 1, 1, 1, 1, 0, 0, 1, 1],
                      np.float32)                                                  # wilson_cowan.rs:85-92
KUR_DEFAULT = np.array([1.0], np.float32)                                          # kuramoto_model.rs:53-55
RWW_DEFAULT = np.array([0.270, 0.108, 154, 0.641, 100, 0.6, 0.2609, 0.33], np.float32)  # rww.rs:61-63
LINEAR_DEFAULT = np.array([-10.0], np.float32)                                      # linear.rs:33-35
SUP_HOPF_DEFAULT = np.array([-0.5, 1.0], np.float32)                                # sup_hopf.rs:35-37
HOPFIELD_DEFAULT = np.array([1.0, 5.0, 0.0], np.float32)                            # hopfield.rs:36-38
CB2D_DEFAULT = np.array([1.0, -4.0, 1.0, 2.0], np.float32)                          # coombes_byrne2d.rs:41-43
EPI_CODIM3_DEFAULT = np.array([-0.02285, 0.3448, 0.2014, -0.07465, 0.3351, 0.2053,
                               1.0, 0.4, 0.001, 0.3, 0.0, 1.0, 1.0], np.float32)    # epileptor_codim3.rs:51-57
EPI_CODIM3_SLOWMOD_DEFAULT = np.array([
    -0.02285, 0.3448, 0.2014, -0.07465, 0.3351, 0.2053, 1.0, 0.4, 0.002, 0.3, 0.0, 1.0, 1.0,
    0.68281186, -0.13736244, 0.717565, 0.60745907, 0.11524223, 0.785947,
    0.6668986, -0.2839006, -0.6889461, 0.79066, 0.00755125, -0.6122089, 0.0001, 0.00012],
    np.float32)                                                                    # epileptor_codim3_slowmod.rs:72-83
EPI_RS_DEFAULT = np.array([3.1, 0.45, -1.6, 1.0, 3.0, 1.0, 5.0, 0.00035, 0.0, 10.0, 6.0, 2.0, 0.0, 0.0, 0.0, 1.0,
                           0.0, 1.0, 0.0, -2.0, -10.0, 0.02, 3.0, 1.0, 1.0, 1.0, 1.0, 1.0],
                          np.float32)                                              # epileptor_rs.rs:72-79
ZERLAUT_DEFAULT = np.array([
    10.0, -65.0, -65.0, 200.0, 60.0, 4.0, 0.0, 0.0, 500.0, 1.0, 0.0, -80.0, 1.5, 5.0, 5.0, 5.0, 10000.0,
    0.05, 0.05, 0.2, 400.0, 0.0, 20.0, 0.0, 0.0, 0.0, 0.0, 5.0, 10.5, 1.0,
    -0.04983106, 0.005063551, -0.023470122, 0.0022951514, -0.00041053028,
    0.010547051, -0.03659253, 0.0074374876, 0.0012650647, -0.040721614,
    -0.05149122, 0.004003689, -0.008352013, 0.0002414238, -0.0005070645,
    0.0014345394, -0.01468669, 0.004502706, 0.0028472191, -0.0153578045],
    np.float32)                                        # zerlaut_first.rs:108-128 (zerlaut_second.rs: the same)
KIONEX_DEFAULT = np.array([0.0, 5.5, 0.1, 0.0, 1.0, -40.0, 0.5, -20.0, -0.5, -31.0, 1.0, 4.0, 0.04, 0.001],
                          np.float32)                                              # kionex.rs:58-64
COOMBES_BYRNE_DEFAULT = np.array([0.5, 0.95, -10.0, 1.0, 20.0], np.float32)               # coombes_byrne.rs:46-48
GAST_SD_DEFAULT = np.array([1.0, 10.0, 0.5, 0.0, 2.0, 21.2132, -6.0, 1.0, 0.0], np.float32)  # gast_schmidt_knosche_sd.rs:52-54
GAST_SF_DEFAULT = np.array([1.0, 10.0, 10.0, 0.0, 2.0, 21.2132, 1.0, 1.0, 0.0], np.float32)  # gast_schmidt_knosche_sf.rs:52-54
LARTER_BREAKSPEAR_DEFAULT = np.array([1.1, 2.0, 0.5, 6.7, 0.7, 1.0, -0.7, -0.5, 0.53, -0.01, 0.3, 0.0,
                                      0.15, 0.15, 0.3, 0.65, 0.7, 2.0, 2.0, 0.4, 1.0, 0.4, 0.1, 0.1, 0.3,
                                      0.25, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0], np.float32)  # larter_breakspear.rs:74-84
EPILEPTOR_2D_DEFAULT = np.array([3.1, -1.6, 1.0, 3.0, 0.0, 1.0, 5.0, 0.00035, 0.0, 0.0, 1.0, 0.0],
                                np.float32)                                       # epileptor2d.rs:47-49
EPILEPTOR_DEFAULT = np.array([3.1, 0.45, -1.6, 1.0, 3.0, 1.0, 5.0, 0.00035, 0.0, 10.0, 6.0, 2.0, 0.0, 0.0,
                              0.0, 1.0, 0.0], np.float32)                         # epileptor.rs:57-62
RWW_EXC_INH_DEFAULT = np.array([310.0, 125.0, 0.160, 0.000641, 100.0, 1.4, 0.15, 1.0, 615.0, 177.0, 0.087,
                                0.001, 10.0, 1.0, 0.7, 0.382, 0.0, 2.0, 0.0], np.float32)  # rww_exc_inh.rs:61-68
DECO_BALANCED_DEFAULT = np.append(RWW_EXC_INH_DEFAULT, np.float32(1.0))          # deco_balanced_exc_inh.rs:62-69
ZETTERBERG_JANSEN_DEFAULT = np.array([3.25, 22.0, 0.1, 0.05, 0.0025, 6.0, 0.56, 135.0, 108.0, 33.75, 33.75,
                                      15.0, 1.0, 1.0, 1.0, 0.12, 0.12, 0.12], np.float32)  # zetterberg_jansen.rs:67-75
REDUCED_FHN_DEFAULT = np.array([3.0, 0.45, 0.9, 0.5, 0.15, 0.15, 0.35, 0.0], np.float32)  # reduced_fhn.rs:45-47
REDUCED_HR_DEFAULT = np.array([0.006, 1.0, 3.0, 1.0, 5.0, 4.0, -1.6, 0.5, 0.1, 0.15, 0.3, 3.3],
                              np.float32)                                         # reduced_hr.rs:52-54
DUMONT_GUTKIN_DEFAULT = np.array([0.0, 1.0, -5.0, 10.0, 0.0, 1.0, -5.0, 10.0, 1.0, 0.0, 10.0, 0.0, 15.0, 5.0],
                                 np.float32)                                      # dumont_gutkin.rs:67-73
DEFAULTS = {L.G2DO: G2DO_DEFAULT, L.MPR: MPR_DEFAULT, L.JANSEN_RIT: JR_DEFAULT,
            L.WILSON_COWAN: WC_DEFAULT, L.KURAMOTO: KUR_DEFAULT, L.RWW: RWW_DEFAULT}
# defaults of the models that run on the multi-subnetwork path only
MULTI_DEFAULTS = {L.LINEAR: LINEAR_DEFAULT, L.SUP_HOPF: SUP_HOPF_DEFAULT, L.HOPFIELD: HOPFIELD_DEFAULT,
                  L.COOMBES_BYRNE_2D: CB2D_DEFAULT, L.EPILEPTOR_CODIM3: EPI_CODIM3_DEFAULT,
                  L.EPILEPTOR_CODIM3_SLOWMOD: EPI_CODIM3_SLOWMOD_DEFAULT, L.EPILEPTOR_RS: EPI_RS_DEFAULT,
                  L.ZERLAUT_FIRST: ZERLAUT_DEFAULT, L.ZERLAUT_SECOND: ZERLAUT_DEFAULT, L.KIONEX: KIONEX_DEFAULT,
                  L.COOMBES_BYRNE: COOMBES_BYRNE_DEFAULT, L.GAST_SD: GAST_SD_DEFAULT, L.GAST_SF: GAST_SF_DEFAULT,
                  L.LARTER_BREAKSPEAR: LARTER_BREAKSPEAR_DEFAULT, L.EPILEPTOR_2D: EPILEPTOR_2D_DEFAULT,
                  L.EPILEPTOR: EPILEPTOR_DEFAULT, L.RWW_EXC_INH: RWW_EXC_INH_DEFAULT,
                  L.DECO_BALANCED: DECO_BALANCED_DEFAULT, L.ZETTERBERG_JANSEN: ZETTERBERG_JANSEN_DEFAULT,
                  L.REDUCED_FHN: REDUCED_FHN_DEFAULT, L.REDUCED_HR: REDUCED_HR_DEFAULT,
                  L.DUMONT_GUTKIN: DUMONT_GUTKIN_DEFAULT}


def dense_to_csr(W: np.ndarray):
    """CSR by target row (row = target), column order within rows."""
    W = np.asarray(W, np.float32)
    rows, cols = np.nonzero(W)
    indptr = np.zeros(W.shape[0] + 1, np.int32)
    np.add.at(indptr, rows + 1, 1)
    indptr = np.cumsum(indptr).astype(np.int32)
    return indptr, cols.astype(np.int32), W[rows, cols].astype(np.float32)


def lognormal_connectome(rng, n: int, density: float, symmetric: bool = True):
    """Lognormal(0,1) weights on a random `density` fraction of the off-diagonal
    pairs (upper triangle mirrored when symmetric), zero diagonal, max = 1."""
    W = np.zeros((n, n), np.float64)
    if symmetric:
        iu = np.triu_indices(n, 1)
        m = len(iu[0])
        k = int(round(density * m))
        sel = rng.choice(m, size=k, replace=False)
        w = rng.lognormal(0.0, 1.0, size=k)
        W[iu[0][sel], iu[1][sel]] = w
        W = W + W.T
    else:
        off = np.flatnonzero(~np.eye(n, dtype=bool))
        k = int(round(density * len(off)))
        sel = rng.choice(off, size=k, replace=False)
        W.flat[sel] = rng.lognormal(0.0, 1.0, size=k)
    if W.max() > 0:
        W /= W.max()
    return W.astype(np.float32)


def symmetric_uniform(rng, n: int, lo: float, hi: float) -> np.ndarray:
    L_ = rng.uniform(lo, hi, size=(n, n))
    L_ = np.triu(L_, 1)
    return (L_ + L_.T).astype(np.float32)


def delays_from_lengths(lengths: np.ndarray, speed: float, dt: float) -> np.ndarray:
    """Integer delays round(L / speed / dt): the stated rule for these workloads
    (the reference's resolve_delays uses ceil and is never called, config.rs:436-452)."""
    return np.rint(np.asarray(lengths, np.float64) / speed / dt).astype(np.int32)


# ---------------------------------------------------------------- workloads --

def demo_g2do_74(seed: int = 1) -> Instance:
    """Plain forward run: G2DO, 74 nodes, Heun dt=0.1, 10k steps, nsig=0, scalar
    self-coupling 0.01 with Linear a=0.01 (HybridEngine semantics: w * I)."""
    rng = np.random.default_rng(seed)
    N = 74
    x0 = rng.normal(0.0, 0.1, size=(N, 2)).astype(np.float32)
    return Instance(model=L.G2DO, n_nodes=N, n_steps=10_000, dt=0.1, params=G2DO_DEFAULT.copy(),
                    x0=x0, integrator=L.HEUN, engine=L.ENGINE_PLAIN, wkind=L.W_SCALAR_DIAG,
                    scalar_w=0.01, cfn=L.CF_LINEAR, cparams=(0.01,), record_traj=True)


def sweep_g2do_74(n_sweep: int = 128, n_steps: int = 10_000, seed: int = 2, noise_seed: int = 7,
            explicit_noise: bool = False) -> list[Instance]:
    """Batched sweep: G2DO, 74 nodes, Linear G log-uniform [1e-3, 1e-1] x param a
    U[-2, 2]; symmetric 20% lognormal connectome; per-edge delays
    round(U[10,150] mm / 3 mm/ms / 0.1 ms) <= 500; HeunStochastic nsig=0.01."""
    rng = np.random.default_rng(seed)
    N, dt = 74, 0.1
    W = lognormal_connectome(np.random.default_rng(seed + 1), N, 0.20)
    Lm = symmetric_uniform(np.random.default_rng(seed + 2), N, 10.0, 150.0)
    D = delays_from_lengths(Lm, 3.0, dt)
    indptr, indices, weights = dense_to_csr(W)
    delays = D[np.repeat(np.arange(N), np.diff(indptr)), indices].astype(np.int32)
    G = (10.0 ** rng.uniform(-3.0, -1.0, size=n_sweep)).astype(np.float32)
    a = rng.uniform(-2.0, 2.0, size=n_sweep).astype(np.float32)
    x0 = np.random.default_rng(seed + 3).normal(0.0, 0.1, size=(N, 2)).astype(np.float32)
    nsig = np.full(2, 0.01, np.float32)
    out = []
    for i in range(n_sweep):
        p = G2DO_DEFAULT.copy()
        p[2] = a[i]
        noise = philox_normal(noise_seed, i, n_steps * N * 2).reshape(n_steps, N, 2) if explicit_noise else None
        out.append(Instance(model=L.G2DO, n_nodes=N, n_steps=n_steps, dt=dt, params=p, x0=x0,
                            integrator=L.HEUN_STOCH, wkind=L.W_CSR, cfn=L.CF_LINEAR,
                            cparams=(float(G[i]), 0.0), indptr=indptr, indices=indices,
                            weights=weights, edge_delays=delays, delay_mode=L.DELAY_PER_EDGE,
                            nsig=nsig, noise=noise, noise_seed=noise_seed, noise_stream=i))
    return out


def ragged_mpr_jr(n_inst: int = 4096, seed: int = 4, t_scale: float = 1.0, bold: bool = True) -> list[Instance]:
    """Ragged batch: MPR (Linear, BOLD TR=720 ms) or JR (legacy SigmoidalJansenRit
    on y0), N in {68,74,84,164,400}, lognormal CSR at U[5%,20%], delays <= 40 ms
    at dt=0.05 (<= 800 steps), sim length U[500,5000] ms (integer steps), RK4."""
    rng = np.random.default_rng(seed)
    dt = 0.05
    sizes = np.array([68, 74, 84, 164, 400])
    nets = {}
    out = []
    for i in range(n_inst):
        model = L.MPR if rng.random() < 0.5 else L.JANSEN_RIT
        N = int(rng.choice(sizes))
        dens = float(rng.uniform(0.05, 0.20))
        key = (N, round(dens, 2))
        if key not in nets:   # connectomes shared by instances of one (N, density) class
            crng = np.random.default_rng([seed, N, int(round(dens * 100))])
            W = lognormal_connectome(crng, N, round(dens, 2), symmetric=False)
            Lm = crng.uniform(10.0, 120.0, size=(N, N)).astype(np.float32)
            ip, ix, w = dense_to_csr(W)
            d = delays_from_lengths(Lm[np.repeat(np.arange(N), np.diff(ip)), ix], 3.0, dt)
            nets[key] = (ip, ix, w, d.astype(np.int32))
        ip, ix, w, d = nets[key]
        T = int(round(rng.uniform(500.0, 5000.0) / dt * t_scale))
        if model == L.MPR:
            p = MPR_DEFAULT.copy()
            p[2] = float(rng.uniform(-6.0, -4.0))
            x0 = np.stack([np.full(N, 0.1), np.full(N, -2.0)], 1) + rng.normal(0, 0.01, (N, 2))
            G = float(10 ** rng.uniform(-3, -1.3))
            out.append(Instance(model=L.MPR, n_nodes=N, n_steps=T, dt=dt, params=p,
                                x0=x0.astype(np.float32), integrator=L.RK4, wkind=L.W_CSR,
                                cfn=L.CF_LINEAR, cparams=(G, 0.0), indptr=ip, indices=ix, weights=w,
                                edge_delays=d, delay_mode=L.DELAY_PER_EDGE,
                                bold_period=10, bold_tr_ms=720.0, record_bold=bold))
        else:
            p = JR_DEFAULT.copy()
            x0 = rng.normal(0.0, 0.01, (N, 6)).astype(np.float32)
            a = float(rng.uniform(0.5, 2.0))
            out.append(Instance(model=L.JANSEN_RIT, n_nodes=N, n_steps=T, dt=dt, params=p, x0=x0,
                                integrator=L.RK4, wkind=L.W_CSR, cfn=L.CF_SJR, cparams=(a,),
                                indptr=ip, indices=ix, weights=w, edge_delays=d,
                                delay_mode=L.DELAY_PER_EDGE))
    return out


def fc_fit_mpr_84(n_inst: int = 256, n_steps: int = 20_000, seed: int = 5, target_seed: int = 6):
    """Differentiation: MPR, 84 nodes, CSR 10% lognormal, per-edge delays
    round(U[10,90] mm / 3 mm/ms / 0.1 ms) <= 300 steps, Heun dt = 0.1 ms,
    per-instance BOLD on the native 1 ms grid (bold_period 10), per-instance
    G and eta perturbations. Returns (instances, hidden instance for the FC
    target): the target is the FC of the hidden-parameter run."""
    rng = np.random.default_rng(seed)
    N, dt = 84, 0.1
    W = lognormal_connectome(np.random.default_rng(seed + 100), N, 0.10, symmetric=False)
    ip, ix, w = dense_to_csr(W)
    lengths = np.random.default_rng(seed + 200).uniform(10.0, 90.0, len(ix))
    d = delays_from_lengths(lengths, 3.0, dt)
    x0 = np.stack([np.full(N, 0.1), np.full(N, -2.0)], 1).astype(np.float32)

    def make(G, eta):
        return Instance(model=L.MPR, n_nodes=N, n_steps=n_steps, dt=dt, params=MPR_DEFAULT.copy(), x0=x0,
                        integrator=L.HEUN, wkind=L.W_CSR, cfn=L.CF_LINEAR, cparams=(float(G), 0.0),
                        indptr=ip, indices=ix, weights=w, edge_delays=d, delay_mode=L.DELAY_PER_EDGE,
                        node_param_idx=2, node_param=eta.astype(np.float32), bold_period=10)
    eta0 = np.full(N, MPR_DEFAULT[2], np.float32)
    insts = [make(0.3 * (1.0 + 0.2 * rng.standard_normal()), eta0 + 0.1 * rng.standard_normal(N))
             for _ in range(n_inst)]
    trng = np.random.default_rng(target_seed)
    hidden = make(0.4, eta0 + 0.2 * trng.standard_normal(N))
    return insts, hidden


def kuramoto_order_1000(n_inst: int = 512, n_steps: int = 10_000, seed: int = 7, max_delay: int = 400):
    """Large-network adjoint: Kuramoto, 1000 nodes placed uniformly in a
    150 mm cube, CSR 5% lognormal, delays round(dist / 3 mm/ms / 0.1 ms)
    clipped to 400 steps (40 ms), natural frequencies N(10 Hz, 1 Hz) in rad/ms,
    RK4 dt = 0.1 ms, G = linspace(0, 2) over the batch. Returns (instances,
    order-parameter target trace [n_steps])."""
    rng = np.random.default_rng(seed)
    N, dt = 1000, 0.1
    pos = rng.uniform(0.0, 150.0, (N, 3))
    W = lognormal_connectome(np.random.default_rng(seed + 100), N, 0.05, symmetric=False)
    ip, ix, w = dense_to_csr(W)
    tgt = np.repeat(np.arange(N), np.diff(ip))
    dist = np.linalg.norm(pos[tgt] - pos[ix], axis=1)
    d = np.minimum(delays_from_lengths(dist, 3.0, dt), max_delay).astype(np.int32)
    omega = (2.0 * np.pi * rng.normal(10.0, 1.0, N) / 1000.0).astype(np.float32)
    x0 = rng.uniform(0.0, 2.0 * np.pi, (N, 1)).astype(np.float32)
    Gs = np.linspace(0.0, 2.0, n_inst)
    insts = [Instance(model=L.KURAMOTO, n_nodes=N, n_steps=n_steps, dt=dt, params=np.array([0.0], np.float32),
                      x0=x0, integrator=L.RK4, wkind=L.W_CSR, cfn=L.CF_KURAMOTO, cparams=(float(G),),
                      indptr=ip, indices=ix, weights=w, edge_delays=d, delay_mode=L.DELAY_PER_EDGE,
                      node_param_idx=0, node_param=omega) for G in Gs]
    t = np.arange(n_steps) * dt
    target = (0.5 + 0.3 * np.sin(2.0 * np.pi * t / 400.0) * np.exp(-t / 2000.0)).astype(np.float32)
    return insts, target


def node_steps(insts) -> float:
    return float(sum(I.n_nodes * I.n_steps for I in insts))

