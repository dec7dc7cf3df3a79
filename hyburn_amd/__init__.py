"""hyburn_amd: MI355X-native (gfx950) batched neural-mass integration.

A from-scratch HIP implementation of hyburn's batched integration path
(Heun/RK4 stepping of six neural-mass models with dense/sparse delayed
coupling and a Balloon-Windkessel BOLD monitor) behind a C ABI
(include/hyburn_hip.h). This package holds the ctypes bindings and the
host-side mirror of the reference's engine API.
"""
from . import _lib
from ._lib import HyburnError
from .instances import Instance, Output, Plan, run_instances

__all__ = ["Instance", "Output", "Plan", "run_instances", "HyburnError", "_lib"]
