"""Pins the CPU oracle against the reference's own known-answer unit tests
(the literals in src/model/*.rs, src/engine/{coupling,sparse,bold,bold_monitor}.rs)
and against closed forms the mathematics fixes."""
import math

import numpy as np
import pytest

from hyburn_amd import _lib as L
from hyburn_amd.synth import G2DO_DEFAULT, JR_DEFAULT, MPR_DEFAULT, RWW_DEFAULT, WC_DEFAULT

pytestmark = []


# ---------------------------------------------------------------- models ----

def test_g2do_equilibrium_defaults(orc):
    # g2do.rs:88-113: dV = 0, dW = -0.04 at (0,0)
    d = orc.dfun(L.G2DO, G2DO_DEFAULT, [0.0, 0.0], [0.0])
    assert abs(d[0]) < 1e-5 and abs(d[1] + 0.04) < 1e-4


def test_g2do_simplified_form(orc):
    # g2do.rs:115-142: V=0.5, W=0.3 -> dV ~ 0.7583, dW = 1.2
    p = np.array([1, 0, 1, 1, 0, 1, 0, 1 / 3, 1, 1, 1, 1], np.float32)
    d = orc.dfun(L.G2DO, p, [0.5, 0.3], [0.0])
    assert abs(d[0] - 0.7583) < 1e-3 and abs(d[1] - 1.2) < 1e-3
    # exact value of the simplified form: W - V^3/3 + V
    assert abs(d[0] - (0.3 - 0.125 / 3 + 0.5)) < 1e-6


def test_g2do_swept_current_enters_like_I(orc):
    # the swept form (c0 + I_s)*gamma equals the plain form with params[1] = I_s
    p = G2DO_DEFAULT.copy()
    p[1] = 0.7
    a = orc.dfun(L.G2DO, p, [0.3, -0.2], [0.05], dtype=np.float64)
    b = orc.dfun(L.G2DO, G2DO_DEFAULT, [0.3, -0.2], [0.05], sweep_form=1, sweep_val=0.7, dtype=np.float64)
    np.testing.assert_allclose(a, b, rtol=1e-7)


def test_mpr_at_zero(orc):
    # mpr.rs:73-95: dr = 1/pi, dV = -5
    d = orc.dfun(L.MPR, MPR_DEFAULT, [0.0, 0.0], [0.0, 0.0])
    assert abs(d[0] - 1 / math.pi) < 1e-4 and abs(d[1] + 5.0) < 1e-4


def test_mpr_with_state(orc):
    # mpr.rs:97-122
    d = orc.dfun(L.MPR, MPR_DEFAULT, [0.5, 0.3], [0.0, 0.0])
    assert abs(d[0] - (1 / math.pi + 0.3)) < 1e-3
    assert abs(d[1] - (0.09 - math.pi ** 2 * 0.25 - 5.0 + 7.5)) < 1e-3


def test_mpr_clamp(orc):
    # mpr.rs:124-135
    x = orc.clamp(L.MPR, [-0.5, 1.0])
    assert x[0] == 0.0 and x[1] == 1.0


def test_jr_at_zero(orc):
    # jansen_rit.rs:92-133
    d = orc.dfun(L.JANSEN_RIT, JR_DEFAULT, np.zeros(6), [0.0])
    nu, r, v0 = 0.0025, 0.56, 5.52
    s0 = 2 * nu / (1 + math.exp(r * v0))
    assert np.all(np.abs(d[:3]) < 1e-6)
    assert abs(d[3] - 3.25 * 0.1 * s0) < 1e-6
    assert abs(d[4] - 3.25 * 0.1 * (0.22 + 0.8 * 135 * s0)) < 1e-5
    assert abs(d[5] - 22 * 0.05 * 0.25 * 135 * s0) < 1e-5


def test_jr_nonzero_state_and_coupling(orc):
    # jansen_rit.rs:135-190
    d = orc.dfun(L.JANSEN_RIT, JR_DEFAULT, [0.1, 0.2, 0.05, 0, 0, 0], [0.1])
    assert np.all(np.abs(d[:3]) < 1e-6)
    assert abs(d[3] + 9.24e-4) < 5e-4
    assert abs(d[4] - 0.2755) < 0.01
    assert abs(d[5] - 0.0428) < 0.005


def test_kuramoto_dfun(orc):
    # kuramoto_model.rs:65-72
    d = orc.dfun(L.KURAMOTO, [1.0], [0.5], [0.3])
    assert abs(d[0] - 1.3) < 1e-5


def test_kuramoto_clamp_wraps(orc):
    # dfun.rs:188-191: x - floor(x / 2pi) * 2pi
    x = orc.clamp(L.KURAMOTO, [7.0])
    assert abs(x[0] - (7.0 - 2 * math.pi)) < 1e-5
    x = orc.clamp(L.KURAMOTO, [-1.0])
    assert abs(x[0] - (2 * math.pi - 1.0)) < 1e-5


def test_rww_at_zero(orc):
    # rww.rs:73-96
    d = orc.dfun(L.RWW, RWW_DEFAULT, [0.0], [0.0])
    x = 0.33
    axb = 0.270 * x - 0.108
    h = axb / (1 - math.exp(-154 * axb))
    assert abs(d[0] - h * 0.641) < 1e-5


def test_rww_clamp(orc):
    # rww.rs:98-118
    assert orc.clamp(L.RWW, [-0.2])[0] == 0.0
    assert orc.clamp(L.RWW, [1.5])[0] == 1.0


def test_wc_at_zero_and_clamp(orc):
    # wilson_cowan.rs:102-132
    d = orc.dfun(L.WILSON_COWAN, WC_DEFAULT, [0.0, 0.0], [0.0])
    assert np.all(np.abs(d) < 1e-5)
    x = orc.clamp(L.WILSON_COWAN, [-0.5, 1.5])
    assert x[0] == 0.0 and x[1] == 1.0


def test_linear_and_sup_hopf_kats(orc):
    # linear.rs:45-55: x = 1, c = 0 -> dx = gamma = -10; sup_hopf.rs:45-55: zero at the origin
    d = orc.dfun(L.LINEAR, [-10.0], np.array([1.0], np.float32), np.zeros(1, np.float32))
    assert abs(d[0] + 10.0) < 1e-4
    d = orc.dfun(L.SUP_HOPF, [-0.5, 1.0], np.zeros(2, np.float32), np.zeros(2, np.float32))
    assert np.all(np.abs(d) < 1e-5)
    # closed forms (dfun.rs:522-557): dx = gamma x + c; the Hopf normal form
    d = orc.dfun(L.LINEAR, [-3.0], np.array([0.5], np.float32), np.array([0.25], np.float32))
    assert d[0] == np.float32(-1.25)
    x, c = np.array([0.3, -0.4], np.float32), np.array([0.1, 0.2], np.float32)
    d = orc.dfun(L.SUP_HOPF, [-0.5, 1.0], x, c)
    np.testing.assert_allclose(d, [0.275, 0.8], rtol=1e-6)
    # without coupling the radial part is (a - r^2) r^2 and the rotation drops out
    d = orc.dfun(L.SUP_HOPF, [0.7, 2.5], x, np.zeros(2, np.float32))
    r2 = float(x[0]) ** 2 + float(x[1]) ** 2
    assert abs(float(x[0] * d[0] + x[1] * d[1]) - (0.7 - r2) * r2) < 1e-6
    for m, p, x in ((L.LINEAR, [-10.0], [0.7]), (L.SUP_HOPF, [-0.5, 1.0], [0.3, -0.2])):
        x = np.array(x, np.float32)
        c = np.full(len(x), 0.05, np.float32)
        np.testing.assert_allclose(orc.dfun(m, p, x, c), orc.dfun(m, p, x, c, dtype=np.float64), rtol=1e-6)


def test_hopfield_kats(orc):
    # hopfield.rs:46-55: static (dynamic = 0), x = 2, c = 1 -> dx = -1
    d = orc.dfun(L.HOPFIELD, [1.0, 5.0, 0.0], np.array([2.0, 0.0], np.float32), np.array([1.0, 0.0], np.float32))
    assert abs(d[0] + 1.0) < 1e-4 and d[1] == 0.0
    # dynamic: dtheta = (-theta + c1) / tauT; static: 0; taux floored at f32 epsilon
    x, c = np.array([0.5, 0.3], np.float32), np.array([0.1, 0.2], np.float32)
    np.testing.assert_allclose(orc.dfun(L.HOPFIELD, [2.0, 4.0, 1.0], x, c), [-0.2, -0.025], rtol=1e-6)
    assert orc.dfun(L.HOPFIELD, [2.0, 4.0, 0.0], x, c)[1] == 0.0
    d = orc.dfun(L.HOPFIELD, [0.0, 4.0, 1.0], x, c)
    assert d[0] == np.float32(np.float32(-0.4) * np.float32(1.0 / np.float32(1.1920929e-07)))


def test_coombes_byrne_2d_kats(orc):
    # coombes_byrne2d.rs:51-64: at zero, dr = Delta / pi, dV = eta
    p = [1.0, -4.0, 1.0, 2.0]
    d = orc.dfun(L.COOMBES_BYRNE_2D, p, np.zeros(2, np.float32), np.zeros(2, np.float32))
    assert abs(d[0] - 1.0 / math.pi) < 1e-4 and abs(d[1] - 2.0) < 1e-4
    # the written formula at a state, in f64 (c_v does not enter)
    r, v, cr = 0.3, -0.7, 0.05
    dl, vs, k, eta = 0.8, -3.0, 1.5, 1.2
    want = [v * 2 * r + dl / math.pi - r * r * k * math.pi + cr,
            -v * v + eta - r * r * math.pi ** 2 + (-v + vs) * k * math.pi * r]
    d = orc.dfun(L.COOMBES_BYRNE_2D, [dl, vs, k, eta], np.array([r, v], np.float32), np.array([cr, 9.0], np.float32))
    np.testing.assert_allclose(d, want, rtol=1e-5)
    # clamp_batch (dfun.rs:190-194): r >= 0, V free
    np.testing.assert_array_equal(orc.clamp(L.COOMBES_BYRNE_2D, [-0.5, -2.0]), [0.0, -2.0])


def test_models_fp32_vs_fp64_twins_agree(orc):
    rng = np.random.default_rng(0)
    cases = [(L.G2DO, G2DO_DEFAULT, 2, 1), (L.MPR, MPR_DEFAULT, 2, 2), (L.JANSEN_RIT, JR_DEFAULT, 6, 1),
             (L.WILSON_COWAN, WC_DEFAULT, 2, 1), (L.KURAMOTO, np.array([1.0], np.float32), 1, 1),
             (L.RWW, RWW_DEFAULT, 1, 1)]
    for m, p, nv, nc in cases:
        for _ in range(20):
            x = rng.uniform(0.05, 0.9, nv)
            c = rng.uniform(-0.1, 0.1, nc)
            a = orc.dfun(m, p, x.astype(np.float32), c.astype(np.float32))
            b = orc.dfun(m, p, x.astype(np.float32).astype(np.float64), c.astype(np.float32).astype(np.float64),
                         dtype=np.float64)
            np.testing.assert_allclose(a, b, rtol=2e-5, atol=1e-6)


# ---------------------------------------------------------------- coupling --

def test_dense_linear_pipeline(orc):
    # coupling.rs:701-725: W=[[1,0,0],[0,1,1]], x=[[1,2],[3,4],[5,6]], a=2, b=1 -> 3,5,17,21
    W = np.array([[1, 0, 0], [0, 1, 1]], np.float32)
    x = np.array([[1, 2], [3, 4], [5, 6]], np.float32)
    out = orc.coupling(L.CF_LINEAR, [2.0, 1.0], x, dense=W)
    np.testing.assert_array_equal(out.ravel(), [3, 5, 17, 21])


def test_kuramoto_dense_classic(orc):
    # coupling.rs:674-699: +-0.25
    W = np.array([[0, 0.5], [0.5, 0]], np.float32)
    th = np.array([[0.0], [math.pi / 2]], np.float32)
    out = orc.coupling(L.CF_KURAMOTO, [1.0], th, x_tgt=th, dense=W, n_src=2)
    assert abs(out[0, 0] - 0.25) < 1e-5 and abs(out[1, 0] + 0.25) < 1e-5


def test_kuramoto_self_coupling_zero(orc):
    # coupling.rs:622-644: identical phases -> 0
    th = np.full((2, 1), 1.23, np.float32)
    W = np.array([[0, 0.5], [0.5, 0]], np.float32)
    out = orc.coupling(L.CF_KURAMOTO, [1.0], th, x_tgt=th, dense=W, n_src=2)
    assert np.all(np.abs(out) < 1e-5)


def test_sigmoidal_pipeline(orc):
    # coupling.rs:727-748: sigmoid(0.5 + 1.5) with cmin=0,cmax=1
    W = np.array([[1.0, 1.0]], np.float32)
    x = np.array([[0.5], [1.5]], np.float32)
    out = orc.coupling(L.CF_SIGMOIDAL, [0.0, 1.0, 0.0, 1.0, 1.0], x, dense=W)
    assert abs(out[0, 0] - 1 / (1 + math.exp(-2.0))) < 1e-5


def test_difference_post_rowsum(orc):
    # coupling.rs:660-672: a * (gx - x_i * rowsum) = 0.1 * (5 - 2*3) = -0.1
    W = np.array([[2.0, 1.0]], np.float32)        # rowsum 3
    x = np.array([[2.0], [1.0]], np.float32)       # W @ x = 5
    out = orc.coupling(L.CF_DIFFERENCE, [0.1], x, x_tgt=np.array([[2.0]], np.float32), dense=W)
    assert abs(out[0, 0] + 0.1) < 1e-6


def test_csr_matvec_basic(orc):
    # sparse.rs:134-154: -> 4, 9, 1
    import ctypes as C
    data = np.array([2, 3, 1], np.float32)
    ind = np.array([1, 2, 0], np.int32)
    ipt = np.array([0, 1, 2, 3], np.int32)
    x = np.array([1, 2, 3], np.float32)
    y = np.zeros(3, np.float32)
    orc.lib().hb_orc_csr_matvec_f32(orc._p(data), orc._p(ind, orc.ip), orc._p(ipt, orc.ip), 3, orc._p(x), orc._p(y))
    np.testing.assert_array_equal(y, [4, 9, 1])


def test_sparse_matches_dense_ring(orc):
    # sparse.rs:156-200 / engine/mod.rs:183-245: 5-node directed ring, w = 0.1
    n = 5
    W = np.zeros((n, n), np.float32)
    for i in range(n):
        W[i, (i - 1) % n] = 0.1
    from hyburn_amd.synth import dense_to_csr
    x = np.arange(1, n + 1, dtype=np.float32).reshape(n, 1)
    a = orc.coupling(L.CF_LINEAR, [1.0, 0.0], x, dense=W)
    b = orc.coupling(L.CF_LINEAR, [1.0, 0.0], x, csr=dense_to_csr(W))
    np.testing.assert_array_equal(a, b)
    np.testing.assert_allclose(a[:, 0], 0.1 * np.roll(x[:, 0], 1), rtol=1e-6)


# ---------------------------------------------------------------- BOLD ------

def _bold_run(orc, z, dt_s, steps, n=2):
    import ctypes as C
    s = np.zeros(n, np.float32); f = np.ones(n, np.float32); v = np.ones(n, np.float32); q = np.ones(n, np.float32)
    zz = np.full(n, z, np.float32)
    sig = np.zeros(n, np.float32)
    out = []
    for _ in range(steps):
        orc.lib().hb_orc_bold_step_f32(n, orc._p(zz), dt_s, orc._p(s), orc._p(f), orc._p(v), orc._p(q))
        orc.lib().hb_orc_bold_signal_f32(n, orc._p(v), orc._p(q), orc._p(sig))
        out.append(sig.copy())
    return np.array(out), (s, f, v, q)


def test_bold_zero_input_fixed_point(orc):
    # bold.rs:174-188 and the fixed point s=0,f=v=q=1 -> signal exactly 0
    sig, (s, f, v, q) = _bold_run(orc, 0.0, 0.001, 100)
    assert np.all(np.isfinite(sig)) and np.all(np.abs(sig[-1]) < 1e-3)
    assert np.all(s == 0) and np.all(f == 1) and np.all(v == 1) and np.all(q == 1)
    assert np.all(sig == 0.0)


def test_bold_ramps_up(orc):
    # bold.rs:190-204: constant input 1 for 5 s at 1 ms -> > 0.1
    sig, _ = _bold_run(orc, 1.0, 0.001, 5000)
    assert sig[-1, 0] > 0.1


def test_bold_negative_input_rectified(orc):
    # z = max(u, 0): negative input is the zero-input fixed point exactly
    sig, _ = _bold_run(orc, -3.0, 0.001, 50)
    assert np.all(sig == 0.0)


def test_bold_monitor_shapes(orc):
    # bold_monitor.rs:263-279: 20 steps, period 5 -> 4 samples; TR = dt_bold -> unchanged
    from hyburn_amd.instances import Instance
    I = Instance(model=L.G2DO, n_nodes=3, n_steps=20, dt=2.0, params=G2DO_DEFAULT, x0=np.zeros((3, 2), np.float32),
                 integrator=L.EULER, bold_period=5, bold_tr_ms=10.0, record_bold=True)
    r = orc.run(I)
    assert r["bold_native"].shape == (4, 3) and r["bold"].shape == (4, 3)
    assert I.bold_rows() == 4


def test_bold_tr_downsample_count(orc):
    # bold_monitor.rs:281-294: dt=1 ms, period 10, TR 25 ms, 100 steps -> floor(100/25) = 4 rows
    from hyburn_amd.instances import Instance
    I = Instance(model=L.G2DO, n_nodes=2, n_steps=100, dt=1.0, params=G2DO_DEFAULT, x0=np.full((2, 2), 0.1, np.float32),
                 integrator=L.EULER, bold_period=10, bold_tr_ms=25.0, record_bold=True)
    r = orc.run(I)
    assert r["bold_native"].shape[0] == 10
    assert r["bold"].shape[0] == 4 == I.bold_rows()
    # t=25 lies between native samples at 20 and 30 ms: halfway interpolation
    nat = r["bold_native"]
    np.testing.assert_allclose(r["bold"][0], nat[1] + (nat[2] - nat[1]) * np.float32(0.5), rtol=1e-6)
    # t=100 is the last native sample exactly -> nearest
    np.testing.assert_array_equal(r["bold"][3], nat[9])


# ---------------------------------------------------------------- FC --------

def test_fc_matches_pearson(orc):
    rng = np.random.default_rng(3)
    X = rng.normal(size=(500, 7)).astype(np.float32)
    X[:, 3] = 2 * X[:, 1] + 0.5                        # perfectly correlated pair
    X[:, 5] = 4.0                                      # constant: std -> 1, FC 0 off-diagonal
    F = orc.fc(X)
    ref = np.corrcoef(X[:, [0, 1, 2, 3, 4, 6]].astype(np.float64), rowvar=False)
    idx = [0, 1, 2, 3, 4, 6]
    np.testing.assert_allclose(F[np.ix_(idx, idx)], ref, atol=1e-6)
    assert abs(F[1, 3] - 1.0) < 1e-6
    assert F[5, 5] == 0.0 and np.all(F[5, np.arange(7) != 5] == 0.0)
    np.testing.assert_array_equal(F, F.T)


def test_fc_reference_kats(orc):
    # fc.rs:225-249: perfectly (anti-)correlated two-node series
    F = orc.fc(np.array([[1, 1], [2, 2], [3, 3]], np.float32))
    np.testing.assert_allclose(F, 1.0, atol=1e-5)
    F = orc.fc(np.array([[1, -1], [2, -2], [3, -3]], np.float32))
    assert abs(F[0, 1] + 1.0) < 1e-5 and abs(F[1, 0] + 1.0) < 1e-5


def test_fcd_reference_kat(orc):
    # fc.rs:268-275: sin(i) over 10 steps x 3 nodes, window 4, stride 2 -> 4 windows
    X = np.sin(np.arange(30, dtype=np.float32)).reshape(10, 3)
    D = orc.fcd(X, 4, 2)
    assert D.shape == (4, 4)
    assert abs(D[0, 0] - 1.0) < 1e-5


def _upper(F):
    i, j = np.triu_indices(F.shape[0], 1)
    return F[i, j].astype(np.float64)


def test_fcd_matches_corrcoef_of_window_fcs(orc):
    # fc.rs:94-179: windows [w*S, w*S+L), FC per window, Pearson of the i<j entries
    rng = np.random.default_rng(11)
    T, N, L, S = 230, 9, 40, 15
    X = np.cumsum(rng.normal(size=(T, N)), axis=0).astype(np.float32)
    D = orc.fcd(X, L, S)
    W = (T - L) // S + 1
    assert D.shape == (W, W)
    ups = np.stack([_upper(np.corrcoef(X[w * S:w * S + L].astype(np.float64), rowvar=False))
                    for w in range(W)])
    np.testing.assert_allclose(D, np.corrcoef(ups), atol=2e-6)
    assert np.all(np.diag(D) == 1.0)
    np.testing.assert_array_equal(D, D.T)


def test_fcd_closed_forms(orc):
    rng = np.random.default_rng(12)
    # a series that repeats every S steps: every window has the same FC -> all ones
    base = rng.normal(size=(20, 6)).astype(np.float32)
    X = np.tile(base, (8, 1))
    D = orc.fcd(X, 40, 20)
    np.testing.assert_allclose(D, 1.0, atol=1e-6)
    # two nodes: one upper entry per window -> zero variance -> 0 off-diagonal
    D2 = orc.fcd(rng.normal(size=(100, 2)).astype(np.float32), 20, 10)
    assert D2.shape == (9, 9)
    np.testing.assert_array_equal(D2, np.eye(9, dtype=np.float32))
    # fewer than two windows -> empty; the asserts of fcd_matrix -> errors
    assert orc.fcd(X, 160, 1).shape == (0, 0)
    assert orc.fcd(X, 100, 100).shape == (0, 0)
    for L, S in [(1, 1), (161, 1), (10, 0)]:
        with pytest.raises(ValueError):
            orc.fcd(X, L, S)


# ---------------------------------------------------------------- Welch ---

def _sine(n, f, fs):
    return np.sin(np.arange(n, dtype=np.float64) * 2.0 * np.pi * f / fs).astype(np.float32)


def test_spectral_reference_kats(orc):
    # spectral.rs:157-264
    F = orc.spectral_features(_sine(1024, 10.0, 100.0), 100.0)
    assert F.shape == (9,) and F[2] > 0.5 and abs(F[5] - 10.0) < 3.0
    F = orc.spectral_features(_sine(1024, 6.0, 100.0), 100.0)
    assert F[1] > 0.5 and F[0] < 0.3
    assert orc.spectral_features(_sine(2048, 40.0, 200.0), 200.0)[4] > 0.5
    assert np.all(np.isnan(orc.spectral_features(np.ones(4, np.float32), 100.0)))
    assert not np.any(np.isinf(orc.spectral_features(_sine(32, 10.0, 100.0), 100.0)))


def test_welch_against_direct_dft(orc):
    # brute-force DFT of every zero-padded segment, on a tiny input
    rng = np.random.default_rng(4)
    y = rng.normal(size=23)
    nfft, fs, w = 20, 7.0, 6
    psd, freqs = orc.welch_psd(y, nfft, fs, w)
    nf, hop, m = 32, 3, y.mean()
    segs = [(y[s:s + w] - m) for s in range(0, 23 - w + 1, hop)]
    ref = np.zeros(nf // 2 + 1)
    for d in segs:
        for k in range(nf // 2 + 1):
            X = sum(d[i] * np.exp(-2j * np.pi * k * i / nf) for i in range(w))
            ref[k] += abs(X) ** 2
    ref /= len(segs) * w * fs
    np.testing.assert_allclose(psd, ref, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(freqs, np.arange(17) * fs / nf)


def test_welch_closed_forms(orc):
    # one segment, window = nfft: Parseval over the completed two-sided spectrum
    rng = np.random.default_rng(5)
    y = rng.normal(size=64)
    psd, _ = orc.welch_psd(y, 64, 1.0, 64)
    two_sided = psd[0] + 2 * psd[1:-1].sum() + psd[-1]
    d = y - y.mean()
    np.testing.assert_allclose(two_sided * 64 * 1.0, 64 * (d * d).sum(), rtol=1e-12)
    # a cosine on bin 5 of a 64-point window: all power there, |X|^2 = (64/2)^2
    c = np.cos(2 * np.pi * 5 * np.arange(64) / 64)
    psd, _ = orc.welch_psd(c, 64, 1.0, 64)
    assert abs(psd[5] * 64 - 32.0 ** 2) < 1e-9 and np.all(np.abs(np.delete(psd, 5)) < 1e-20)
    # a window longer than the series: no segment fits, the PSD is zero
    psd, _ = orc.welch_psd(y[:10], 16, 1.0, 12)
    assert psd.shape == (9,) and np.all(psd == 0.0)
    # constant series: zero power -> NaN features
    assert np.all(np.isnan(orc.spectral_features(np.full(100, 3.0, np.float32), 100.0)))


# ---------------------------------------------------------------- delays ----

def test_slot_tables_ring(orc):
    # runtime.rs:162-204: d==0 or s==0 -> current; d > s -> zero; else s - d
    H = 4
    table = [[orc.source_time(L.DELAY_PER_EDGE, s, d, H) for d in range(H)] for s in range(6)]
    assert table == [[0, 0, 0, 0], [1, 0, -1, -1], [2, 1, 0, -1], [3, 2, 1, 0], [4, 3, 2, 1], [5, 4, 3, 2]]


def test_slot_tables_batch_compat(orc):
    # engine.rs:541-589: the window holds x_{s-H+1..s}; slot (s-d+H)%H is read.
    # The reference's example: H=4, d=3, s=4 reads effective delay 2.
    H = 4
    assert orc.source_time(L.DELAY_BATCH_COMPAT, 4, 3, H) == 2
    for s in range(1, 40):
        for d in range(1, H):
            t = orc.source_time(L.DELAY_BATCH_COMPAT, s, d, H)
            if d > s:
                assert t == -1
                continue
            slot = (s - d + H) % H
            want = s - (H - 1) + slot
            assert t == (want if want >= 0 else -1)
            # the read equals the true delay only when s = H-1 (mod H)
            if s % H == H - 1:
                assert t == s - d
