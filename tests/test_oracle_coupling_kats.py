"""Oracle pins for the coupling functions and the Balloon-Windkessel step that
the model KATs do not reach: HyperbolicTangent, PreSigmoidal, ScaledLinear and
SigmoidalJansenRit (classic and legacy), each against the reference's own
unit-test literals, plus closed forms that a dropped term, a wrong sign or a
swapped operand fails.

Reference tests restated (src/engine/coupling.rs):
  test_tanh_pre 577-587, test_dense_coupling_tanh_pipeline 751-773,
  test_presigmoidal_static_p_neq1_theta_neq0 775-793,
  test_scaled_linear_post 845-855, test_scaled_linear_pipeline 857-871,
  test_scaled_linear_pre_identity 873-883,
  test_sigmoidal_jansen_rit_pre_2cvar 952-963,
  test_sigmoidal_jansen_rit_pre_single_cvar_fallback 965-975,
  test_sigmoidal_jansen_rit_classic_pre 977-1002,
  test_sigmoidal_jansen_rit_classic_post 1004-1024,
  test_sigmoidal_jansen_rit_legacy_post_is_identity 1026-1047,
  parameter parsing test_sjr_from_params_* 1196-1223.
Balloon-Windkessel (src/engine/bold.rs:106-166): the fixed point under a
constant input and the linearised small-signal response, both derived from
the equations independently of the oracle's code.
"""
import math

import numpy as np
import pytest

from hyburn_amd import _lib as L


# ------------------------------------------------------------------ tanh ----

@pytest.mark.parametrize("cpu_form", [False, True])
def test_tanh_pre_at_zero(orc, cpu_form):
    # coupling.rs:577-587: a=b=sigma=1, midpoint=0, x=0 -> a * (1 + tanh 0) = 1
    out = orc.pre(L.CF_TANH, [1.0, 1.0, 0.0, 1.0], np.zeros((1, 1)), cpu_form=cpu_form)
    assert out.shape == (1, 1) and abs(out[0, 0] - 1.0) < 1e-6


def test_tanh_dense_pipeline(orc):
    # coupling.rs:751-773: W=[[1,1]], x=[1,-1] -> (1 + tanh 1) + (1 + tanh(-1)) = 2
    out = orc.coupling(L.CF_TANH, [1.0, 1.0, 0.0, 1.0], np.array([[1.0], [-1.0]]), dense=np.ones((1, 2)))
    expected = (1.0 + math.tanh(1.0)) + (1.0 + math.tanh(-1.0))
    assert abs(out[0, 0] - expected) < 1e-4


def test_tanh_parameters_enter_as_written(orc):
    # closed form a * (1 + tanh((x*b - midpoint) / sigma)): each parameter moved
    # alone moves the value the way the formula says (a swapped b/sigma or a
    # dropped midpoint fails one of these)
    x = np.array([[0.7], [-1.3], [2.2]], np.float32)
    for a, b, mid, sig in [(2.0, 1.0, 0.0, 1.0), (1.0, 3.0, 0.0, 1.0), (1.0, 1.0, 0.5, 1.0), (1.0, 1.0, 0.0, 4.0),
                           (0.3, -2.0, 0.25, 0.5)]:
        got = orc.pre(L.CF_TANH, [a, b, mid, sig], x)[:, 0]
        want = a * (1.0 + np.tanh((x[:, 0].astype(np.float64) * b - mid) / sig))
        np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-7)
    # post is the identity (coupling.rs:342)
    g = np.array([[0.25], [-3.5]], np.float32)
    np.testing.assert_array_equal(orc.post(L.CF_TANH, [2.0, 3.0, 0.5, 4.0], g), g)


# ------------------------------------------------------------ presigmoidal --

@pytest.mark.parametrize("cpu_form", [False, True])
def test_presigmoidal_static(orc, cpu_form):
    # coupling.rs:775-793: h=2, q=1, g=3, p=0.5, theta=2; x = 4..7 (2 rows x 2 cvars)
    x = np.array([[4.0, 5.0], [6.0, 7.0]], np.float32)
    out = orc.pre(L.CF_PRESIGMOIDAL, [2.0, 1.0, 3.0, 0.5, 2.0], x, cpu_form=cpu_form)
    want = 2.0 * (1.0 + np.tanh(3.0 * (0.5 * x.astype(np.float64) - 2.0)))
    np.testing.assert_allclose(out, want, atol=1e-5)
    # x = 4: the inner term is 0, so the value is h * q = 2 exactly (the
    # reference's check against the old formula)
    assert out[0, 0] == 2.0
    # defaults h=q=g=p=1, theta=0.5 (coupling.rs:191-198)
    d = orc.pre(L.CF_PRESIGMOIDAL, [], np.array([[0.5], [1.5]], np.float32))
    np.testing.assert_allclose(d[:, 0], [1.0, 1.0 + math.tanh(1.0)], rtol=1e-6)
    np.testing.assert_array_equal(orc.post(L.CF_PRESIGMOIDAL, [2.0, 1.0, 3.0, 0.5, 2.0], out), out)


# ----------------------------------------------------------- scaled linear --

def test_scaled_linear_post(orc):
    # coupling.rs:845-855: a=3, b=2, gx=5 -> (5 - 2) * 3 = 9
    out = orc.post(L.CF_SCALED_LINEAR, [3.0, 2.0], np.array([[5.0]], np.float32))
    assert out[0, 0] == 9.0


def test_scaled_linear_pipeline_and_pre(orc):
    # coupling.rs:857-871: W = I, x = [3, 5], a=2, b=1 -> [4, 8]
    out = orc.coupling(L.CF_SCALED_LINEAR, [2.0, 1.0], np.array([[3.0], [5.0]]), dense=np.eye(2))
    np.testing.assert_array_equal(out[:, 0], [4.0, 8.0])
    # coupling.rs:873-883: pre is the identity
    x = np.array([[1.0, 2.0], [3.0, 4.0]], np.float32)
    np.testing.assert_array_equal(orc.pre(L.CF_SCALED_LINEAR, [2.0, 1.0], x), x)
    # (gx - b) * a, not gx * a - b or gx * a + b
    assert orc.post(L.CF_SCALED_LINEAR, [2.0, 1.0], np.array([[0.0]], np.float32))[0, 0] == -2.0


# ------------------------------------------------------- sigmoidal JR -------

def test_sjr_classic_pre(orc):
    # coupling.rs:977-1002: a=2, cmin=0, cmax=0.01, r=0.56, midpoint=6;
    # rows (8, 2) and (4, 0): diff = x0 - x1 -> cmin + (cmax - cmin) / (1 + exp(r (mid - diff)))
    x = np.array([[8.0, 2.0], [4.0, 0.0]], np.float32)
    for cpu_form in (False, True):
        out = orc.pre(L.CF_SJR, [2.0, 0.0, 0.01, 0.56, 6.0], x, cpu_form=cpu_form)
        assert out.shape == (2, 1), "classic pre collapses to one channel"
        for row, diff in ((0, 6.0), (1, 4.0)):
            assert abs(out[row, 0] - 0.01 / (1.0 + math.exp(0.56 * (6.0 - diff)))) < 1e-5
    # at diff == midpoint the sigmoid is exactly half way: cmin + (cmax - cmin) / 2
    out = orc.pre(L.CF_SJR, [2.0, 0.001, 0.011, 0.56, 6.0], np.array([[9.0, 3.0]], np.float32))
    assert out[0, 0] == np.float32(0.001) + np.float32(0.01) / np.float32(2.0)


def test_sjr_classic_post(orc):
    # coupling.rs:1004-1024: a=3 -> gx * a: [1, 2] -> [3, 6]
    out = orc.post(L.CF_SJR, [3.0, 0.0, 0.01, 0.56, 6.0], np.array([[1.0], [2.0]], np.float32),
                   x_tgt=np.zeros((2, 1), np.float32))
    np.testing.assert_array_equal(out[:, 0], [3.0, 6.0])


def test_sjr_legacy_post_is_identity(orc):
    # coupling.rs:1026-1047
    g = np.array([[0.01], [0.02]], np.float32)
    out = orc.post(L.CF_SJR, [1.0, 0.005, 0.56, 6.0], g, x_tgt=np.zeros((2, 1), np.float32))
    np.testing.assert_array_equal(out, g)


@pytest.mark.parametrize("cpu_form", [False, True])
def test_sjr_legacy_pre_closed_forms(orc, cpu_form):
    """Legacy pre(x) = a * 2 * e0 / (1 + exp(-r (x0 - x1 - v0))) (coupling.rs:234-236,
    310-313) with x1 = 0 for one cvar (coupling.rs:223-227)."""
    a, e0, r, v0 = 1.7, 0.005, 0.56, 6.0
    cp = [a, e0, r, v0]
    # at x = v0 the denominator is exactly 2: the value is a * e0 up to the f32
    # product (a * 2) * e0 and one exact halving
    at_v0 = orc.pre(L.CF_SJR, cp, np.array([[v0]], np.float32), cpu_form=cpu_form)[0, 0]
    assert at_v0 == np.float32(np.float32(np.float32(a) * np.float32(2.0)) * np.float32(e0)) / np.float32(2.0)
    # off the midpoint, against the FP64 closed form (a wrong sign of r or of
    # v0 moves these by orders of magnitude more than the tolerance)
    xs = np.array([[v0 + 5.0], [v0 - 5.0], [0.0], [20.0]], np.float32)
    got = orc.pre(L.CF_SJR, cp, xs, cpu_form=cpu_form)[:, 0]
    want = a * 2.0 * e0 / (1.0 + np.exp(-r * (xs[:, 0].astype(np.float64) - v0)))
    np.testing.assert_allclose(got, want, rtol=3e-6)
    # two cvars: the difference x0 - x1 (coupling.rs:952-963: rows (6, 4) are
    # finite and >= 0; here also the closed form)
    x2 = np.array([[6.0, 4.0], [6.0, 4.0]], np.float32)
    got2 = orc.pre(L.CF_SJR, cp, x2, cpu_form=cpu_form)[:, 0]
    np.testing.assert_allclose(got2, a * 2.0 * e0 / (1.0 + math.exp(-r * (2.0 - v0))), rtol=3e-6)
    assert np.all(np.isfinite(got2)) and np.all(got2 >= 0)
    # coupling.rs:965-975: one cvar rows (6), (8) are finite
    assert np.all(np.isfinite(orc.pre(L.CF_SJR, cp, np.array([[6.0], [8.0]], np.float32))))


def test_sjr_parameter_parsing(orc):
    # coupling.rs:1196-1223 and 156-189: fewer than 5 parameters is legacy
    # [a, e0, r, v0] (defaults e0=0.005, r=0.56, v0=6); 5 or more is classic
    # [a, cmin, cmax, r, midpoint]
    x = np.array([[6.0]], np.float32)
    legacy_defaults = orc.pre(L.CF_SJR, [5.0], x)[0, 0]
    legacy_explicit = orc.pre(L.CF_SJR, [5.0, 0.005, 0.56, 6.0], x)[0, 0]
    assert legacy_defaults == legacy_explicit == np.float32(np.float32(10.0) * np.float32(0.005)) / np.float32(2.0)
    classic = orc.pre(L.CF_SJR, [2.0, 0.0, 0.01, 0.56, 6.0], x)[0, 0]
    assert abs(classic - 0.005) < 1e-9     # cmin + (cmax - cmin) / 2 at diff == midpoint


def test_pre_channel_counts(orc):
    # coupling.rs:805-813: Kuramoto 2 per cvar, SigmoidalJansenRit 1, the rest 1 per cvar
    x1, x2 = np.zeros((3, 1), np.float32), np.zeros((3, 2), np.float32)
    assert orc.pre(L.CF_KURAMOTO, [1.0], x1).shape == (3, 2)
    assert orc.pre(L.CF_SJR, [2.0, 0.0, 0.01, 0.56, 6.0], x2).shape == (3, 1)
    assert orc.pre(L.CF_SJR, [1.0], x1).shape == (3, 1)
    for cfn in (L.CF_LINEAR, L.CF_DIFFERENCE, L.CF_SIGMOIDAL, L.CF_SCALED_LINEAR, L.CF_TANH, L.CF_PRESIGMOIDAL):
        assert orc.pre(cfn, [], x2).shape == (3, 2)


# ------------------------------------------------------ Balloon-Windkessel --

# the reference holds the BW parameters as f32 (bold.rs:62-76)
E0, ALPHA, TAU_S, TAU_F, TAU_0 = (float(np.float32(c)) for c in (0.34, 0.32, 1.54, 0.50, 1.02))
V0, K1, K2, K3 = 4.0, 7.0, 2.0, 2.0


def _bw64(orc, z, dt, n, state=None):
    import ctypes as C
    N = len(z)
    zz = np.ascontiguousarray(z, np.float64)
    s, f, v, q = (np.zeros(N), np.ones(N), np.ones(N), np.ones(N)) if state is None else [a.copy() for a in state]
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    for _ in range(n):
        orc.lib().hb_orc_bold_step_f64(N, P(zz), dt, P(s), P(f), P(v), P(q))
    return s, f, v, q


def _signal(v, q):
    return V0 * (K1 * (1 - q) + K2 * (1 - q / v) + K3 * (1 - v))


def test_bold_fixed_point_under_constant_input(orc):
    """With a constant input z the BW equations rest where ds = df = dv = dq = 0:
    s = 0, f = 1 + z * tau_f (from ds), v = f^alpha (from dv) and
    q = v * (1 - (1 - E0)^(1/f)) / E0 (from dq). Started there, the FP64
    oracle stays there; a dropped or sign-flipped term in any of the four
    right-hand sides moves it. Started from rest, s oscillates into its
    clamp at 0, so df = s >= 0 ratchets f past 1 + z * tau_f; it then rests at
    s = 0 with v and q on the same dv = dq = 0 relations of its own f."""
    import ctypes as C
    z = np.array([0.0, 0.05, 0.2, 0.6])
    f_star = 1.0 + z * TAU_F
    v_star = f_star ** ALPHA
    q_star = v_star * (1.0 - (1.0 - E0) ** (1.0 / f_star)) / E0
    s, f, v, q = _bw64(orc, z, 0.01, 1000, (np.zeros(4), f_star, v_star, q_star))
    np.testing.assert_allclose(s, 0.0, atol=1e-13)
    np.testing.assert_allclose(f, f_star, rtol=1e-12)
    np.testing.assert_allclose(v, v_star, rtol=1e-12)
    np.testing.assert_allclose(q, q_star, rtol=1e-12)
    v32, q32 = v.astype(np.float32), q.astype(np.float32)
    s32 = np.zeros(len(z), np.float32)
    orc.lib().hb_orc_bold_signal_f32(len(z), orc._p(v32), orc._p(q32), orc._p(s32))
    np.testing.assert_allclose(s32, _signal(v_star, q_star), rtol=1e-5, atol=1e-6)
    # from rest: 200 s, then at rest on the ratcheted f
    s, f, v, q = _bw64(orc, z, 0.01, 20000)
    np.testing.assert_allclose(s, 0.0, atol=1e-12)
    assert np.all(f >= f_star - 1e-12) and f[0] == 1.0 and np.all(f[1:] > f_star[1:])
    np.testing.assert_allclose(v, f ** ALPHA, rtol=1e-9)
    np.testing.assert_allclose(q, v * (1.0 - (1.0 - E0) ** (1.0 / f)) / E0, rtol=1e-9)


def test_bold_linearised_small_signal_response(orc):
    """Around s = 0, f = v = q = 1 a small input z drives the linear system
    ds = z - s/tau_s - (f - 1)/tau_f, df = s, dv = (df_ - dv_/alpha)/tau_0,
    dq = (g1 df_ - dq_ - (1/alpha - 1) dv_)/tau_0 with
    g1 = (E0 + (1 - E0) ln(1 - E0)) / E0 (the derivative of
    f (1 - (1 - E0)^(1/f)) / E0 at f = 1). Forward Euler of that system
    matches the oracle's nonlinear step to O(z) relative; the time constants,
    the oxygen-extraction derivative and every sign enter."""
    z, dt, n = 1e-5, 0.01, 200    # 2 s: s stays > 0, clear of its clamp
    g1 = (E0 + (1.0 - E0) * math.log(1.0 - E0)) / E0
    A = np.array([[-1.0 / TAU_S, -1.0 / TAU_F, 0.0, 0.0],
                  [1.0, 0.0, 0.0, 0.0],
                  [0.0, 1.0 / TAU_0, -1.0 / (ALPHA * TAU_0), 0.0],
                  [0.0, g1 / TAU_0, -(1.0 / ALPHA - 1.0) / TAU_0, -1.0 / TAU_0]])
    b = np.array([1.0, 0.0, 0.0, 0.0])
    d = np.zeros(4)
    lin = []
    for _ in range(n):
        d = d + dt * (A @ d + b * z)
        lin.append(d.copy())
    lin = np.array(lin)
    s, f, v, q = np.zeros(1), np.ones(1), np.ones(1), np.ones(1)
    got = []
    for _ in range(n):
        s, f, v, q = _bw64(orc, np.array([z]), dt, 1, (s, f, v, q))
        got.append([s[0], f[0] - 1.0, v[0] - 1.0, q[0] - 1.0])
    got = np.array(got)
    scale = np.max(np.abs(lin), axis=0)
    assert np.all(scale > 1e-7)                    # every variable responds
    assert np.max(np.abs(got - lin) / scale) < 1e-4
    # the signal's linearisation: V0 (-(k1 + k2) dq + (k2 - k3) dv)
    sig_lin = V0 * (-(K1 + K2) * lin[-1, 3] + (K2 - K3) * lin[-1, 2])
    sig_got = _signal(1.0 + got[-1, 2], 1.0 + got[-1, 3])
    assert abs(sig_got - sig_lin) <= 1e-4 * abs(sig_lin)
