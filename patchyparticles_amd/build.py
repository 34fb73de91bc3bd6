"""Build the gfx950 library ``libpp_hip.so`` (and the C host programs) in-tree.

    python -m patchyparticles_amd.build        # or __graft_entry__.build()

hipcc flags:
  --offload-arch=gfx950          MI355X only, one code path
  -ffp-contract=off              no fused multiply-add: the pair test must round
                                 where the reference's C source rounds
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                 keep the compiler from emitting packed FP32 ops
  -mllvm -amdgpu-sched-strategy=max-ilp   (the kernel files)
                                 schedule for instruction-level parallelism: the
                                 colour passes' wavefronts run mostly alone on
                                 their SIMDs, bound by each trial's dependent
                                 chain (C2 +0.8 %, C5 +1.9 %, the learned pass
                                 at C4 +1 % measured)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(CSRC, "host")
LIB = os.path.join(PKG, "libpp_hip.so")
BIN = os.path.join(PKG, "bin")

HIP_SOURCES = ["pp_kernels.hip", "pp_learn.hip", "pp_api.cpp", "pp_net.cpp", "pp_host.cpp", "pp_noalloc.cpp"]
EXTRA_FLAGS = {src: ["-mllvm", "-amdgpu-sched-strategy=max-ilp"] for src in ("pp_kernels.hip", "pp_learn.hip")}
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise FileNotFoundError("hipcc not found")


def _run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def build_lib(verbose=True):
    objs = []
    obj_dir = os.path.join(PKG, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    for src in HIP_SOURCES:
        obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        _run([hipcc(), *HIPCC_FLAGS, *EXTRA_FLAGS.get(src, []), *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_host():
    """C host programs: PatchyParticles_hip (input file / .rrr / outputs driver)
    and generator_hip, both linked to libpp_hip.so next to them."""
    if not os.path.exists(os.path.join(HOST, "pp_main.c")):
        return None
    os.makedirs(BIN, exist_ok=True)
    flags = ["-O2", "-std=gnu99", "-Wall", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"), "-I", HOST]
    link = ["-L", PKG, "-lpp_hip", "-Wl,-rpath,$ORIGIN/..", "-lm"]
    common = [os.path.join(HOST, f) for f in ("pp_input.c", "pp_rrr.c", "pp_output.c")]
    out = os.path.join(BIN, "PatchyParticles_hip")
    _run(["gcc", *flags, os.path.join(HOST, "pp_main.c"), *common, *link, "-o", out])
    _run(["gcc", *flags, os.path.join(HOST, "pp_generator.c"), *link, "-o", os.path.join(BIN, "generator_hip")])
    return out


def build_checks():
    """GPU checkers run by the tests: kf_filter_check (the colour pass's
    filtered pair test against the reference arithmetic at the thresholds),
    devmath_check (the device's sin, cos, exp, log and sqrt against the
    host's C library) and philox_check (the device's Philox blocks, draw
    counters and colour orders, compared with the oracle by the test),
    compiled with the library's flags."""
    built = []
    flags = [f for f in HIPCC_FLAGS if f != "-fPIC"]
    for name in ("kf_filter_check", "devmath_check", "philox_check"):
        src = os.path.join(ROOT, "tests", "gpu_checks", name + ".hip")
        if not os.path.exists(src):
            continue
        os.makedirs(BIN, exist_ok=True)
        out = os.path.join(BIN, name)
        _run([hipcc(), *flags, "-I", CSRC, "-I", os.path.join(ROOT, "include"), src, "-o", out])
        built.append(out)
    return built


def build_all():
    build_lib()
    build_host()
    build_checks()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
