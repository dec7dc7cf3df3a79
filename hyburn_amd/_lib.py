"""ctypes bindings for libhyburn_hip.so (include/hyburn_hip.h).

The library is loaded from this package directory only; importing the
bindings when it has not been built raises immediately. There is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhyburn_hip.so")

# ---- constants (mirror include/hyburn_hip.h) ----
HB_OK, HB_INVALID_CONFIG, HB_INVALID_STATE, HB_BACKEND_ERROR, HB_IO_ERROR = 0, 1, 2, 3, 4
G2DO, MPR, JANSEN_RIT, WILSON_COWAN, KURAMOTO, RWW = 0, 1, 2, 3, 4, 5
LINEAR, SUP_HOPF, HOPFIELD, COOMBES_BYRNE_2D = 6, 7, 8, 9   # on the multi-subnetwork path only
# the models the reference pins with trace fixtures (tests/model_reference_test.rs); multi-subnetwork path only
EPILEPTOR_CODIM3, EPILEPTOR_CODIM3_SLOWMOD, EPILEPTOR_RS, ZERLAUT_FIRST, ZERLAUT_SECOND, KIONEX = 10, 11, 12, 13, 14, 15
# the remaining models of the registry (config.rs:43-72); multi-subnetwork path only
(COOMBES_BYRNE, GAST_SD, GAST_SF, LARTER_BREAKSPEAR, EPILEPTOR_2D, EPILEPTOR, RWW_EXC_INH, DECO_BALANCED,
 ZETTERBERG_JANSEN, REDUCED_FHN, REDUCED_HR, DUMONT_GUTKIN) = range(16, 28)
EULER, HEUN, RK4, EULER_STOCH, HEUN_STOCH, RK4_STOCH = 0, 1, 2, 3, 4, 5
CF_LINEAR, CF_DIFFERENCE, CF_KURAMOTO, CF_SIGMOIDAL, CF_SCALED_LINEAR, CF_TANH, CF_SJR, CF_PRESIGMOIDAL = range(8)
W_NONE, W_SCALAR_MEAN, W_SCALAR_DIAG, W_DENSE, W_CSR = 0, 1, 2, 3, 4
DELAY_PER_EDGE, DELAY_SINGLE, DELAY_BATCH_COMPAT = 0, 1, 2
ENGINE_BATCH, ENGINE_PLAIN = 0, 1
FLAG_TRAJECTORY, FLAG_EXACT, FLAG_NO_GRAPH, FLAG_NO_BLOCKING, FLAG_AUTOTUNE = 0x1, 0x2, 0x4, 0x8, 0x10
FLAG_SEPARATE_BOLD = 0x20
FLAG_KEEP_NOISE, FLAG_RESTART, FLAG_BOLD_BATCH_MEAN, FLAG_NO_MFMA = 0x40, 0x80, 0x100, 0x200
LOSS_FC, LOSS_BOLD_MSE, LOSS_ORDER = 0, 1, 2

MODEL_NAMES = {
    "Generic2dOscillator": G2DO, "MontbrioPazoRoxin": MPR, "JansenRit": JANSEN_RIT,
    "WilsonCowan": WILSON_COWAN, "Kuramoto": KURAMOTO, "ReducedWongWang": RWW,
    "Linear": LINEAR, "SupHopf": SUP_HOPF, "Hopfield": HOPFIELD, "CoombesByrne2D": COOMBES_BYRNE_2D,
    "EpileptorCodim3": EPILEPTOR_CODIM3, "EpileptorCodim3SlowMod": EPILEPTOR_CODIM3_SLOWMOD,
    "EpileptorRestingState": EPILEPTOR_RS, "ZerlautAdaptationFirstOrder": ZERLAUT_FIRST,
    "ZerlautAdaptationSecondOrder": ZERLAUT_SECOND, "KIonEx": KIONEX,
    "CoombesByrne": COOMBES_BYRNE, "GastSchmidtKnoscheSD": GAST_SD, "GastSchmidtKnoscheSF": GAST_SF,
    "LarterBreakspear": LARTER_BREAKSPEAR, "Epileptor2D": EPILEPTOR_2D, "Epileptor": EPILEPTOR,
    "ReducedWongWangExcInh": RWW_EXC_INH, "DecoBalancedExcInh": DECO_BALANCED,
    "ZetterbergJansen": ZETTERBERG_JANSEN, "ReducedSetFitzHughNagumo": REDUCED_FHN,
    "ReducedSetHindmarshRose": REDUCED_HR, "DumontGutkin": DUMONT_GUTKIN,
}
# (nvar, ncvar, nparams), config.rs:43-50
MODEL_INFO = {G2DO: (2, 1, 12), MPR: (2, 2, 7), JANSEN_RIT: (6, 1, 13),
              WILSON_COWAN: (2, 1, 22), KURAMOTO: (1, 1, 1), RWW: (1, 1, 8),
              LINEAR: (1, 1, 1), SUP_HOPF: (2, 2, 2), HOPFIELD: (2, 2, 3),
              COOMBES_BYRNE_2D: (2, 2, 4), EPILEPTOR_CODIM3: (3, 1, 13), EPILEPTOR_CODIM3_SLOWMOD: (5, 1, 27),
              EPILEPTOR_RS: (8, 3, 28), ZERLAUT_FIRST: (5, 1, 50), ZERLAUT_SECOND: (8, 1, 50), KIONEX: (5, 1, 14),
              COOMBES_BYRNE: (4, 4, 5), GAST_SD: (4, 4, 9), GAST_SF: (4, 4, 9), LARTER_BREAKSPEAR: (3, 1, 32),
              EPILEPTOR_2D: (2, 1, 12), EPILEPTOR: (6, 2, 17), RWW_EXC_INH: (2, 1, 19), DECO_BALANCED: (2, 1, 20),
              ZETTERBERG_JANSEN: (12, 1, 18), REDUCED_FHN: (4, 2, 8), REDUCED_HR: (6, 2, 12),
              DUMONT_GUTKIN: (8, 4, 14)}
MULTI_ONLY = (LINEAR, SUP_HOPF, HOPFIELD, COOMBES_BYRNE_2D, EPILEPTOR_CODIM3, EPILEPTOR_CODIM3_SLOWMOD, EPILEPTOR_RS,
              ZERLAUT_FIRST, ZERLAUT_SECOND, KIONEX) + tuple(range(16, 28))
# hybrid_integrator: deterministic Heun for these models (model_prefers_heun, dfun.rs:1871-1873)
PREFERS_HEUN = (G2DO, EPILEPTOR_RS, EPILEPTOR_CODIM3, EPILEPTOR_CODIM3_SLOWMOD, KIONEX, EPILEPTOR, EPILEPTOR_2D)
# the batch-native sweep index per model, dfun.rs:49-72
SWEEP_INDEX = {G2DO: 1, MPR: 4, JANSEN_RIT: 12, WILSON_COWAN: 18, KURAMOTO: 0, RWW: 7}
SWEEP_INDEX.update({m: -1 for m in MULTI_ONLY})
COUPLING_NAMES = {
    "Linear": CF_LINEAR, "Difference": CF_DIFFERENCE, "Kuramoto": CF_KURAMOTO,
    "Sigmoidal": CF_SIGMOIDAL, "ScaledLinear": CF_SCALED_LINEAR, "HyperbolicTangent": CF_TANH,
    "SigmoidalJansenRit": CF_SJR, "PreSigmoidal": CF_PRESIGMOIDAL,
}
INTEGRATOR_NAMES = {
    "euler": EULER, "heun": HEUN, "rk4": RK4, "euler_stochastic": EULER_STOCH,
    "heun_stochastic": HEUN_STOCH, "rk4_stochastic": RK4_STOCH,
}

fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int32)


class HbInstance(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("integrator", C.c_int32), ("engine", C.c_int32),
        ("n_nodes", C.c_int32), ("n_steps", C.c_int32),
        ("dt", C.c_float), ("dt_ms", C.c_double),
        ("params", fp), ("sweep_form", C.c_int32), ("sweep_val", C.c_float),
        ("node_param_idx", C.c_int32), ("node_param", fp),
        ("x0", fp), ("nsig", fp), ("noise", fp),
        ("noise_seed", C.c_uint64), ("noise_stream", C.c_uint64),
        ("wkind", C.c_int32), ("cfn", C.c_int32), ("n_cparams", C.c_int32),
        ("cparams", C.c_float * 8), ("scalar_w", C.c_float),
        ("dense", fp), ("indptr", ip), ("indices", ip), ("weights", fp), ("edge_delays", ip),
        ("nnz", C.c_int32), ("delay_mode", C.c_int32), ("single_delay", C.c_int32),
        ("max_delay_hint", C.c_int32),
        ("bold_period", C.c_int32), ("bold_tr_ms", C.c_double),
        ("tavg", fp), ("final_state", fp), ("traj", fp), ("bold", fp),
        ("bold_cap", C.c_int32), ("bold_n", ip),
        ("step0", C.c_int64), ("history", fp), ("bold_state", fp)]


class HbRunOpts(C.Structure):
    _fields_ = [("device", C.c_int32), ("flags", C.c_uint32), ("block_size", C.c_int32),
                ("grid_waves", C.c_int32)]


class HbRunStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double), ("node_steps", C.c_double),
                ("n_launches", C.c_int32), ("block_size", C.c_int32), ("lds_bytes", C.c_int32)]


class HbSubnetwork(C.Structure):
    _fields_ = [("model", C.c_int32), ("n_nodes", C.c_int32), ("n_params", C.c_int32),
                ("params", fp), ("x0", fp)]


class HbProjection(C.Structure):
    _fields_ = [("wkind", C.c_int32), ("scalar_w", C.c_float), ("dense", fp),
                ("indptr", ip), ("indices", ip), ("data", fp), ("nnz", C.c_int32),
                ("delays", ip), ("n_delays", C.c_int32), ("cfn", C.c_int32),
                ("n_cparams", C.c_int32), ("cparams", C.c_float * 8)]


class HbSimConfig(C.Structure):
    _fields_ = [("sim_length", C.c_double), ("dt", C.c_double), ("integrator", C.c_int32),
                ("n_nsig", C.c_int32), ("nsig", fp),
                ("subnetworks", C.POINTER(HbSubnetwork)), ("n_subnetworks", C.c_int32),
                ("projections", C.POINTER(HbProjection)), ("n_projections", C.c_int32),
                ("bold_period", C.c_int32), ("bold_tr_s", C.c_double),
                ("engine", C.c_int32), ("delay_mode", C.c_int32),
                ("noise_seed", C.c_uint64), ("noise_shared", C.c_int32)]


class HbSweepPoint(C.Structure):
    _fields_ = [("kind", C.c_int32), ("a", C.c_int32), ("b", C.c_int32),
                ("values", fp), ("n_values", C.c_int32)]


class HbResult(C.Structure):
    _fields_ = [("n_sweep", C.c_int32), ("n_nodes", C.c_int32), ("nvar", C.c_int32),
                ("n_steps", C.c_int32), ("tavg", fp), ("final_states", fp),
                ("trajectories", fp), ("bold", fp), ("n_bold", C.c_int32),
                ("elapsed_ms", C.c_double), ("kernel_ms", C.c_double), ("bold_series", C.c_int32),
                ("step", C.c_int64)]


class HbEngineState(C.Structure):
    _fields_ = [("step", C.c_int64), ("dt", C.c_double), ("integrator", C.c_int32), ("n_state", C.c_int32),
                ("n_nodes", C.c_int32), ("nvar", C.c_int32), ("hist_floats", C.c_int64)]


class HbAdjointSpec(C.Structure):
    _fields_ = [("loss_kind", C.c_int32), ("target", fp), ("grad_param", C.c_int32),
                ("loss", fp), ("grad_coupling", fp), ("grad_param_out", fp), ("grad_x0", fp),
                ("final_state", fp)]


class HbAdjointOpts(C.Structure):
    _fields_ = [("device", C.c_int32), ("checkpoint_every", C.c_int32), ("flags", C.c_uint32)]


class HbAdjointStats(C.Structure):
    _fields_ = [("forward_ms", C.c_double), ("backward_ms", C.c_double), ("node_steps", C.c_double),
                ("device_bytes", C.c_int64)]


class HbMultiProjection(C.Structure):
    _fields_ = [("src", C.c_int32), ("tgt", C.c_int32), ("wkind", C.c_int32), ("scalar_w", C.c_float),
                ("dense_w", fp), ("cfn", C.c_int32), ("n_cparams", C.c_int32), ("cparams", C.c_float * 8),
                ("n_map", C.c_int32), ("map_src", C.c_int32 * 4), ("map_tgt", C.c_int32 * 4), ("delay", C.c_int32),
                ("indptr", ip), ("indices", ip), ("csr_w", fp), ("edge_delays", ip), ("nnz", C.c_int32)]


class HbStimulus(C.Structure):
    _fields_ = [("target", C.c_int32), ("pattern", C.c_int32), ("n_params", C.c_int32), ("params", C.c_float * 4)]


class HbMultiSpec(C.Structure):
    _fields_ = [("n_subnets", C.c_int32), ("subnets", C.POINTER(HbSubnetwork)), ("n_projections", C.c_int32),
                ("projections", C.POINTER(HbMultiProjection)), ("dt", C.c_float), ("integrator", C.c_int32),
                ("hybrid", C.c_int32), ("n_sweep", C.c_int32), ("n_steps", C.c_int32), ("sweep_sub", C.c_int32),
                ("sweep_param", C.c_int32), ("sweep_values", fp), ("max_delay", C.c_int32), ("n_nsig", C.c_int32),
                ("nsig", fp), ("noise", fp), ("device", C.c_int32), ("tavg", fp),
                ("final_state", fp), ("kernel_ms", C.c_double), ("step0", C.c_int64), ("state_in", fp),
                ("hist_in", fp), ("hist_out", fp), ("plain", C.c_int32), ("n_stimuli", C.c_int32),
                ("stimuli", C.POINTER(HbStimulus)), ("dt_ms", C.c_double), ("trajectory", fp),
                ("bold_period", C.c_int32), ("bold_tr_ms", C.c_double), ("bold", fp),
                ("point_mode", C.c_int32 * 4), ("point_values", fp), ("point_params", fp),
                ("bold_mean", C.c_int32), ("bold_state", fp)]


MAX_PARAMS = 64   # HB_MAX_PARAMS


# Every exported function of include/hyburn_hip.h with its prototype.
PROTOTYPES = {
    "hb_engine_create": (C.c_int, [C.POINTER(HbSimConfig), C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "hb_engine_destroy": (None, [C.c_void_p]),
    "hb_run_sweep": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, fp, C.c_int32, C.c_int64, C.c_uint32,
                               C.POINTER(HbResult)]),
    "hb_run_sweep_multi": (C.c_int, [C.c_void_p, C.POINTER(HbSweepPoint), C.c_int32, C.c_int64, C.c_uint32,
                                     C.POINTER(HbResult)]),
    "hb_result_free": (None, [C.POINTER(HbResult)]),
    "hb_engine_get_state": (C.c_int, [C.c_void_p, C.POINTER(HbEngineState), fp, fp, fp]),
    "hb_engine_set_state": (C.c_int, [C.c_void_p, C.POINTER(HbEngineState), fp, C.c_int32, fp, fp]),
    "hb_plan_create": (C.c_int, [C.POINTER(HbInstance), C.c_int32, C.POINTER(HbRunOpts), C.POINTER(C.c_void_p)]),
    "hb_plan_run": (C.c_int, [C.c_void_p, C.POINTER(HbRunStats)]),
    "hb_plan_download": (C.c_int, [C.c_void_p]),
    "hb_plan_destroy": (None, [C.c_void_p]),
    "hb_run_instances": (C.c_int, [C.POINTER(HbInstance), C.c_int32, C.POINTER(HbRunOpts), C.POINTER(HbRunStats)]),
    "hb_philox_normal": (C.c_int, [C.c_uint64, C.c_uint64, C.c_int64, fp, C.c_int32]),
    "hb_history_floats": (C.c_int64, [C.POINTER(HbInstance)]),
    "hb_last_error": (C.c_char_p, []),
    "hb_version": (C.c_char_p, []),
    "hb_device_count": (C.c_int, []),
    "hb_device_synchronize": (C.c_int, [C.c_int32]),
    "hb_select_backend": (C.c_char_p, [C.c_char_p]),
    "hb_adjoint_plan_create": (C.c_int, [C.POINTER(HbInstance), C.POINTER(HbAdjointSpec), C.c_int32,
                                         C.POINTER(HbAdjointOpts), C.POINTER(C.c_void_p)]),
    "hb_adjoint_forward": (C.c_int, [C.c_void_p, C.POINTER(HbAdjointStats)]),
    "hb_adjoint_backward": (C.c_int, [C.c_void_p, C.POINTER(HbAdjointStats)]),
    "hb_adjoint_download": (C.c_int, [C.c_void_p]),
    "hb_adjoint_plan_destroy": (None, [C.c_void_p]),
    "hb_autotune_entries": (C.c_int, [ip, C.c_int32]),
    "hb_autotune_clear": (None, []),
    "hb_fc_batch": (C.c_int, [fp, C.c_int32, C.c_int32, C.c_int32, fp, C.c_int32]),
    "hb_fc_batch_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_int32,
                                      C.c_void_p]),
    "hb_fcd_windows": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32]),
    "hb_fcd_batch": (C.c_int, [fp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, fp, C.c_int32]),
    "hb_run_multi": (C.c_int, [C.POINTER(HbMultiSpec)]),
    "hb_bold_rows": (C.c_int32, [C.c_int32, C.c_int32, C.c_double, C.c_double]),
    "hb_bold_downsample": (C.c_int32, [C.POINTER(C.c_double), C.c_int32, fp, C.c_int32, C.c_double, C.c_double, fp,
                                       C.c_int32]),
    "hb_welch_bins": (C.c_int32, [C.c_int32, C.c_int32]),
    "hb_welch_psd_batch": (C.c_int, [fp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32,
                                     C.POINTER(C.c_double), C.c_int32]),
    "hb_spectral_features_batch": (C.c_int, [fp, C.c_int32, C.c_int32, C.c_int32, C.c_double, fp, C.c_int32]),
}

_lib = None


class HyburnError(RuntimeError):
    """A non-zero status from the C ABI (codes mirror src/error.rs:7-23)."""

    NAMES = {1: "InvalidConfig", 2: "InvalidState", 3: "BackendError", 4: "IOError"}

    def __init__(self, code: int, msg: str):
        super().__init__(f"{self.NAMES.get(code, code)}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load libhyburn_hip.so from this package (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run hyburn_amd.build.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != HB_OK:
        raise HyburnError(rc, lib().hb_last_error().decode(errors="replace"))
