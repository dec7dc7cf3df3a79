"""Pins the oracle's restatement of the six models the reference checks with
trace fixtures (EpileptorCodim3, EpileptorCodim3SlowMod, EpileptorRestingState,
ZerlautAdaptationFirstOrder / SecondOrder, KIonEx; dfun.rs:952-1870):

* the reference's own trace tests (tests/model_reference_test.rs:90-156): two
  uncoupled nodes from the same initial state, the plain engine's Heun or
  Euler, the final state against the last row of the TVB fixture, with the
  test's tolerance (relative rtol where |expected| >= 1e-4, else absolute
  1e-4). The final rows are committed in tests/golden/model_reference_final.json
  (written by tests/golden/make_model_reference_final.py);
* the reference's unit-test derivative values (epileptor_rs.rs:88-104,
  zerlaut_first.rs:153-168, zerlaut_second.rs:138-153, kionex.rs:73-90);
* closed forms: the Codim3 fixed point xs is a root of x^3 - mu2 x - mu1 (so
  the fast subsystem is at rest there, and the slow variable moves by
  c (dstar - 0.1 (z - 0.5)^7 - Ks c0), which pins the oracle's xs, made of
  the reference's acos/atan approximations, against the exact root);
* the FP32 twin against the FP64 twin over the fixture runs.
"""
import json
import math
import os

import numpy as np
import pytest

from hyburn_amd import _lib as L
from hyburn_amd.synth import (EPI_CODIM3_DEFAULT, EPI_CODIM3_SLOWMOD_DEFAULT, EPI_RS_DEFAULT, KIONEX_DEFAULT,
                              ZERLAUT_DEFAULT)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "model_reference_final.json")

# model_reference_test.rs:90-156: (model, params, var-major initial state, dt, integrator, steps, fixture, rtol)
TRACE_CASES = {
    "EpileptorCodim3": (L.EPILEPTOR_CODIM3, EPI_CODIM3_DEFAULT, [0.5, 0.5, 0.0, 0.0, 0.1, 0.1], 0.1, L.HEUN, 100,
                        "epileptor_codim3_ref.npy", 0.01),
    "EpileptorCodim3SlowMod": (L.EPILEPTOR_CODIM3_SLOWMOD, EPI_CODIM3_SLOWMOD_DEFAULT,
                               [0.5, 0.5, 0.0, 0.0, 0.05, 0.05, 0.0, 0.0, 0.0, 0.0], 0.1, L.HEUN, 100,
                               "epileptor_codim3_slowmod_ref.npy", 0.01),
    "EpileptorRestingState": (L.EPILEPTOR_RS, EPI_RS_DEFAULT,
                              [-1.6, -1.6, -12.5, -12.5, 3.8, 3.8, -1.0, -1.0, 0.005, 0.005, 0.0, 0.0, 1.0, 1.0,
                               0.0, 0.0], 0.1, L.HEUN, 100, "epileptor_rs_ref.npy", 0.05),
    "ZerlautAdaptationFirstOrder": (L.ZERLAUT_FIRST, ZERLAUT_DEFAULT,
                                    [0.01, 0.01, 0.01, 0.01, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0], 0.1, L.EULER, 100,
                                    "zerlaut_first_ref.npy", 0.10),
    "ZerlautAdaptationSecondOrder": (L.ZERLAUT_SECOND, ZERLAUT_DEFAULT,
                                     [0.01, 0.01, 0.01, 0.01, 0.0001, 0.0001, 0.0, 0.0, 0.0001, 0.0001, 0.0, 0.0,
                                      0.0, 0.0, 0.0, 0.0], 0.1, L.EULER, 100, "zerlaut_second_ref.npy", 0.05),
    "KIonEx": (L.KIONEX, KIONEX_DEFAULT, [0.1, 0.1, -65.0, -65.0, 0.5, 0.5, -0.5, -0.5, -2.0, -2.0], 0.01, L.HEUN,
               1000, "kionex_ref.npy", 0.10),
}


def trace_run(orc, name, dtype=np.float32):
    """Final state of the reference's trace test, variable-major like its fixture rows."""
    m, p, ic, dt, integ, T, _, _ = TRACE_CASES[name]
    nv = L.MODEL_INFO[m][0]
    x0 = np.array(ic, np.float32).reshape(nv, 2).T.copy()          # var-major -> node-major
    _, fin = orc.run_multi([(m, 2, p, x0)], [], dt, T, integ, False, plain=True, dtype=dtype)
    return fin[0][0].T.reshape(-1)


def _reference_allclose(actual, expected, rtol, atol=1e-4):
    """model_reference_test.rs:25-40."""
    for i, (a, e) in enumerate(zip(actual, expected)):
        assert math.isfinite(a), (i, a)
        if abs(e) < atol:
            assert abs(a - e) < atol, (i, a, e)
        else:
            assert abs(a - e) / abs(e) <= rtol, (i, a, e, abs(a - e) / abs(e))


@pytest.mark.parametrize("name", list(TRACE_CASES))
def test_trace_fixture_final_state(orc, name):
    golden = json.load(open(GOLDEN))
    fx = golden[TRACE_CASES[name][6]]
    got = trace_run(orc, name)
    assert len(got) == fx["shape"][1]
    _reference_allclose(got, fx["final_row"], TRACE_CASES[name][7])


@pytest.mark.parametrize("name", list(TRACE_CASES))
def test_fp32_twin_against_fp64(orc, name):
    a = trace_run(orc, name).astype(np.float64)
    b = trace_run(orc, name, dtype=np.float64)
    scale = np.maximum(np.abs(b), 1.0)
    assert np.max(np.abs(a - b) / scale) < 1e-4
    assert not np.array_equal(a, b)   # two arithmetics, not one


def test_reference_unit_test_derivatives(orc):
    # epileptor_rs.rs:88-104 (5%, absolute below 1e-4)
    d = orc.dfun(L.EPILEPTOR_RS, EPI_RS_DEFAULT, np.array([-1.6, -12.5, 3.8, -1.0, 0.005, 0.0, 1.0, 0.0], np.float32),
                 np.zeros(3, np.float32))
    exp = [-1.424, 0.7, -0.00133, 0.355, -0.0005, -0.0016, 0.04, -0.24]
    for a, e in zip(d, exp):
        err = abs(a - e) / abs(e) if abs(e) > 1e-4 else abs(a - e)
        assert err < 0.05, (a, e)
    # zerlaut_first.rs:153-168 and zerlaut_second.rs:138-153 (1%)
    d = orc.dfun(L.ZERLAUT_FIRST, ZERLAUT_DEFAULT, np.array([0.01, 0.01, 0.0, 0.0, 0.0], np.float32),
                 np.zeros(1, np.float32))
    assert abs(d[0] - 0.00365182) / 0.00365182 < 0.01 and abs(d[1] - 0.00501603) / 0.00501603 < 0.01, d
    d = orc.dfun(L.ZERLAUT_SECOND, ZERLAUT_DEFAULT,
                 np.array([0.01, 0.01, 0.0001, 0.0, 0.0001, 0.0, 0.0, 0.0], np.float32), np.zeros(1, np.float32))
    assert np.all(np.isfinite(d)) and abs(d[0] - 0.00365182) / 0.00365182 < 0.01, d
    # kionex.rs:73-90 (5%, absolute below 0.01)
    d = orc.dfun(L.KIONEX, KIONEX_DEFAULT, np.array([0.1, -65.0, 0.5, -0.5, -2.0], np.float32),
                 np.zeros(1, np.float32))
    exp = [-1.50015915, -359.477, -0.107, -0.00509, 0.0012]
    for a, e in zip(d, exp):
        err = abs(a - e) / abs(e) if abs(e) > 0.01 else abs(a - e)
        assert err < 0.05, (a, e)
    # epileptor_codim3(_slowmod).rs: finite derivatives
    for m, p, x in ((L.EPILEPTOR_CODIM3, EPI_CODIM3_DEFAULT, [0.1, 0.1, 0.0]),
                    (L.EPILEPTOR_CODIM3_SLOWMOD, EPI_CODIM3_SLOWMOD_DEFAULT, [0.1, 0.1, 0.0, 0.0, 0.0])):
        assert np.all(np.isfinite(orc.dfun(m, p, np.array(x, np.float32), np.zeros(1, np.float32))))


def _codim3_mus(p, z):
    """mu2, mu1, nu of the great circle through the start and stop points (dfun.rs:958-994), in f64."""
    st = np.array([p[1], -p[0], p[2]], np.float64)
    sp = np.array([p[4], -p[3], p[5]], np.float64)
    a, b = st / np.linalg.norm(st), sp / np.linalg.norm(sp)
    cr = np.cross(a, b)
    f = np.cross(cr, a)
    f /= np.linalg.norm(f)
    v = (math.cos(z) * a + math.sin(z) * f) * p[7]
    return v[0], -v[1], v[2]


@pytest.mark.parametrize("z", [0.0, 0.3, 1.1])
@pytest.mark.parametrize("branch", [1.0, 2.0, 3.0])
def test_codim3_fixed_point_is_a_root(orc, z, branch):
    """At (x, y) = (xs, 0) the fast subsystem rests: dy = xs^3 - mu2 xs - mu1 = 0,
    and dz = c (dstar - 0.1 (z - 0.5)^7 - Ks c0) (dist = 0, modification on). xs is found by bisection on the
    cubic in f64, on the branch the parameters select (trigonometric roots
    ordered as the reference orders them, or the single Cardano root)."""
    p = EPI_CODIM3_DEFAULT.copy()
    p[11] = branch
    p[10] = 0.7     # Ks
    p[8] = 1.0      # c: dz then carries the distance to the oracle's xs unscaled
    mu2, mu1, nu = _codim3_mus(p.astype(np.float64), z)
    disc = mu1 * mu1 / 4 - mu2 ** 3 / 27
    if disc < 0:
        r = math.sqrt(mu2 / 3)
        th = math.acos(max(-1.0, min(1.0, (mu1 / 2) / (mu2 / 3 * r)))) / 3
        shift = {1.0: 0.0, 2.0: -2 * math.pi / 3, 3.0: 2 * math.pi / 3}[branch]
        xs = 2 * r * math.cos(th + shift)
    else:
        xs = np.cbrt(mu1 / 2 + math.sqrt(disc)) + np.cbrt(mu1 / 2 - math.sqrt(disc))
    assert abs(xs ** 3 - mu2 * xs - mu1) < 1e-9
    c0 = 0.25
    d = orc.dfun(L.EPILEPTOR_CODIM3, p, np.array([xs, 0.0, z]), np.array([c0]), dtype=np.float64)
    assert abs(d[0]) == 0.0
    assert abs(d[1]) < 1e-9, d
    # dz = c (dstar - dist - mod - Ks c0): the oracle's xs (acos_approx / atan_approx) is the root to 1e-6
    np.testing.assert_allclose(d[2], p[8] * (p[9] - 0.1 * (z - 0.5) ** 7 - p[10] * c0), rtol=0, atol=1e-6)


def test_clamps(orc):
    assert list(orc.clamp(L.ZERLAUT_FIRST, np.array([-0.5, 1.5, -3, 4, -5], np.float32))) == [0, 1, -3, 4, -5]
    assert list(orc.clamp(L.ZERLAUT_SECOND, np.array([2, -1, -1, -1, -1, -1, -1, -1], np.float32)))[:3] == [1, 0, -1]
    assert list(orc.clamp(L.KIONEX, np.array([-0.1, -65, -1, -1, -1], np.float32))) == [0, -65, -1, -1, -1]
    x = np.array([-1, -2, -3], np.float32)
    assert list(orc.clamp(L.EPILEPTOR_CODIM3, x)) == [-1, -2, -3]
