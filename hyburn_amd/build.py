"""In-tree build of libhyburn_hip.so for gfx950 (hipcc, no JIT cache).

The forward kernel translation unit is compiled twice: with FMA contraction
(the default product path) and with ``-ffp-contract=off`` (the exact path,
selected per run with HB_FLAG_EXACT).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhyburn_hip.so")

ARCH = os.environ.get("HB_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# Packed-FP32 instructions are disabled for all device code.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall"]

# hb_forward.hip is compiled once per kernel unit (one per model and pre
# channel count) and mode, so the units build in parallel.
FORWARD_UNITS = 7
# hb_adjoint.hip: one unit per model (0..5) and one for the shared kernels (6).
ADJOINT_UNITS = 7
UNITS = [
    # (source, object, extra flags, device?)
    *[("hb_forward.hip", f"hb_forward_fast_u{u}.o", ["-DHB_EXACT=0", f"-DHB_UNIT={u}"], True)
      for u in range(FORWARD_UNITS)],
    *[("hb_forward.hip", f"hb_forward_exact_u{u}.o", ["-DHB_EXACT=1", f"-DHB_UNIT={u}", "-ffp-contract=off"], True)
      for u in range(FORWARD_UNITS)],
    ("hb_forward_dispatch.cpp", "hb_forward_dispatch.o", [], False),
    ("hb_noise.hip", "hb_noise.o", [], True),
    ("hb_features.hip", "hb_features.o", [], True),
    ("hb_multi.hip", "hb_multi.o", [], True),
    ("hb_dense.hip", "hb_dense.o", [], True),
    *[("hb_adjoint.hip", f"hb_adjoint_u{u}.o", [f"-DHB_ADJ_UNIT={u}"], True) for u in range(ADJOINT_UNITS)],
    ("hb_host.cpp", "hb_host.o", [], False),
    ("hb_engine.cpp", "hb_engine.o", [], False),
    ("hb_adjoint_host.cpp", "hb_adjoint_host.o", [], False),
]


def _compile(src: str, obj: str, extra: list[str], device: bool) -> tuple[str, int, str]:
    cmd = [HIPCC] + (DEVICE_FLAGS if device else HOST_FLAGS) + extra + [
        "-c", os.path.join(CSRC, src), "-o", os.path.join(BUILD, obj)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass does not know the packed-fp32 feature; that warning is harmless
    err = "\n".join(l for l in (p.stdout + p.stderr).splitlines() if "packed-fp32-ops" not in l)
    return obj, p.returncode, err


def _includes(path: str, seen: set) -> set:
    """Headers of this build that path includes, transitively ("..." includes
    found in csrc/ or include/)."""
    import re
    try:
        text = open(path).read()
    except OSError:
        return seen
    for name in re.findall(r'^\s*#\s*include\s+"([^"]+)"', text, flags=re.M):
        for d in (CSRC, os.path.join(HERE, "..", "include")):
            h = os.path.normpath(os.path.join(d, name))
            if os.path.exists(h) and h not in seen:
                seen.add(h)
                _includes(h, seen)
                break
    return seen


def _stale(src: str, obj: str) -> bool:
    """An object is stale when its source or a header it includes is newer than it."""
    o = os.path.join(BUILD, obj)
    if not os.path.exists(o):
        return True
    s = os.path.join(CSRC, src)
    deps = [s] + sorted(_includes(s, set()))
    return max(os.path.getmtime(d) for d in deps) > os.path.getmtime(o)


def _source_digest() -> str:
    """Digest of every source and header the units read."""
    import hashlib
    h = hashlib.sha256()
    for d in (CSRC, os.path.join(HERE, "..", "include")):
        for f in sorted(os.listdir(d)):
            if f.endswith((".h", ".hip", ".cpp")):
                h.update(f.encode())
                with open(os.path.join(d, f), "rb") as fh:
                    h.update(fh.read())
    return h.hexdigest()


def build(verbose: bool = False, jobs: int | None = None, incremental: bool = False) -> str:
    """Compile every unit for gfx950 and link libhyburn_hip.so in-tree
    (incremental: only units whose source or headers changed since their object).
    A source edited while the units compile would leave objects whose host and
    device passes saw different code (kernel stubs with no device kernel), so
    such a build is refused."""
    os.makedirs(BUILD, exist_ok=True)
    digest0 = _source_digest()
    units = [u for u in UNITS if os.path.exists(os.path.join(CSRC, u[0]))]
    todo = [u for u in units if _stale(u[0], u[1])] if incremental else units
    jobs = jobs or min(os.cpu_count() or 8, 16, max(1, len(todo)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda u: _compile(*u), todo))
    failed = [r for r in results if r[1] != 0]
    for obj, rc, err in results:
        if err.strip() and (verbose or rc != 0):
            print(f"[build] {obj}:\n{err}", file=sys.stderr)
    if failed:
        raise RuntimeError("hipcc failed for: " + ", ".join(r[0] for r in failed))
    if _source_digest() != digest0:
        for u in todo:   # the objects of this build are not trustworthy
            try:
                os.remove(os.path.join(BUILD, u[1]))
            except OSError:
                pass
        raise RuntimeError("sources changed while they compiled; build again")
    objs = [os.path.join(BUILD, u[1]) for u in units]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="-i" in sys.argv))
