"""Builds the native parts in-tree for gfx950.

  libsphk.so                  HIP kernels + C ABI (include/sphk.h)
  libsph_engine.so            C++ SPHEngine / ParticleSystem drop-in over the C ABI
  sph<ext-suffix>.so          pybind11 module `sph` (same names as the reference module)

Usage: python build.py [--only sphk|engine|module]
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 are emitted anywhere in the library.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA results in architectural VGPRs: the surrogate's edge pass adds every result into
    # VGPR accumulators, which would otherwise cost one v_accvgpr_read per value
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wno-unused-command-line-argument",
]

SPHK_SOURCES = ["csrc/sphk_kernels.hip", "csrc/sphk_walk.hip", "csrc/sphk_api.hip",
                "csrc/sphk_surrogate.hip", "csrc/sphk_surrogate_api.hip"]
ENGINE_SOURCES = ["host/particle.cpp", "host/sph_engine.cpp", "host/kernels.cpp", "host/export.cpp"]


def _run(cmd, cwd=HERE):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")
    return out


def build_sphk():
    out = os.path.join(HERE, "libsphk.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", out + ".tmp", *SPHK_SOURCES])
    os.replace(out + ".tmp", out)
    return out


def build_engine():
    out = os.path.join(HERE, "libsph_engine.so")
    srcs = [s for s in ENGINE_SOURCES if os.path.exists(os.path.join(HERE, s))]
    if not srcs:
        return None
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", REPO + "/include",
          "-I", os.path.join(HERE, "host"), "-o", out + ".tmp", *srcs,
          "-L", HERE, "-lsphk", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def build_module():
    src = os.path.join(HERE, "bindings", "sph_module.cpp")
    if not os.path.exists(src):
        return None
    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "sph" + suffix)
    inc = sysconfig.get_paths()["include"]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-fvisibility=hidden",
          "-I", pybind11.get_include(), "-I", inc, "-I", REPO + "/include",
          "-I", os.path.join(HERE, "host"), "-o", out + ".tmp", src,
          "-L", HERE, "-lsph_engine", "-lsphk", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def build_all():
    built = [build_sphk(), build_engine(), build_module()]
    return [b for b in built if b]


if __name__ == "__main__":
    only = sys.argv[sys.argv.index("--only") + 1] if "--only" in sys.argv else None
    fn = {"sphk": build_sphk, "engine": build_engine, "module": build_module}.get(only, build_all)
    print(fn())
