"""Build the native libraries in-tree.

    python -m proteinsim2015_amd.build            # product + oracle (+ reference harness)

Products (git-ignored, shipped with the tree to the GPU host):
  proteinsim2015_amd/libpsim_gpu.so        HIP kernels (gfx950) + C ABI
  proteinsim2015_amd/plugins/*.so          PSim::IForce plugin adaptors and the
                                           PSim::IIntegrator adaptor (GBIntegrator.so)
  proteinsim2015_amd/cli/psim_sim          the simulation program (settings.cfg
                                           in, trial directory out)
Test infrastructure built alongside (never loaded by the product):
  oracle/libpsim_oracle.so                 C restatement of the reference
  oracle/_ref/*                            the reference's own sources, when
                                           the reference checkout is present
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libpsim_gpu.so")
PLUGIN_DIR = os.path.join(PKG, "plugins")
REFERENCE = "/root/reference"

ARCH = os.environ.get("PSIM_OFFLOAD_ARCH", "gfx950")

# HIP device code: gfx950 only, packed-FP32 instructions disabled, no FMA
# contraction unless written explicitly (the integrator must round like the
# reference).
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wshadow", "-Wno-unused-result", "-Wno-unused-function",
]
# tile / launch-shape overrides for sweeps, e.g. PSIM_HIP_DEFINES="PSIM_DENSE_WAVES=6"
HIP_FLAGS += [f"-D{d}" for d in os.environ.get("PSIM_HIP_DEFINES", "").split()]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_product(verbose: bool = False) -> str:
    hipcc = _hipcc()
    tmp = os.path.join(PKG, "_obj")
    os.makedirs(tmp, exist_ok=True)
    jobs = []
    for src in ("psim_engine.hip", "psim_ensemble.hip", "psim_analysis.hip",
                "psim_adjoint.hip", "psim_state.hip", "psim_noalloc.hip",
                "psim_integrator.hip"):
        obj = os.path.join(tmp, src + ".o")
        jobs.append((obj, [hipcc, *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c",
                           os.path.join(CSRC, src), "-o", obj]))
    for src in ("psim_host.cpp", "psim_io.cpp"):
        obj = os.path.join(tmp, src + ".o")
        jobs.append((obj, ["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-I",
                           INCLUDE, "-c", os.path.join(CSRC, src), "-o", obj]))
    # the translation units are independent: compile them side by side
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for f in [ex.submit(_run, cmd, verbose) for _, cmd in jobs]:
            f.result()
    objs = [obj for obj, _ in jobs]
    out = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out, *objs, "-lgomp"],
         verbose)
    os.replace(out, LIB)
    build_plugins(verbose)
    return LIB


PLUGINS = {"LJ": 2, "AO": 4, "Cal": 1}   # file name (as the reference's) -> module bit


def build_plugins(verbose: bool = False) -> None:
    """PSim::IForce plugins (host C++, linked to libpsim_gpu.so)."""
    src = os.path.join(PLUGIN_DIR, "psim_force_plugin.cpp")
    for name, mask in PLUGINS.items():
        out = os.path.join(PLUGIN_DIR, f"{name}.so")
        _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-I", INCLUDE,
              f"-DPSIM_PLUGIN_MASK={mask}", src, "-o", out + ".tmp", "-L", PKG, "-lpsim_gpu",
              "-Wl,-rpath,$ORIGIN/.."], verbose)
        os.replace(out + ".tmp", out)
    integ = os.path.join(PLUGIN_DIR, "GBIntegrator.so")
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-fopenmp", "-Wall", "-I", INCLUDE,
          os.path.join(PLUGIN_DIR, "psim_integrator_plugin.cpp"), "-o", integ + ".tmp", "-L",
          PKG, "-lpsim_gpu", "-Wl,-rpath,$ORIGIN/.."], verbose)
    os.replace(integ + ".tmp", integ)
    demo = os.path.join(PLUGIN_DIR, "psim_host_demo")
    _run(["g++", "-std=c++17", "-O2", "-fopenmp", "-Wall", "-I", INCLUDE,
          os.path.join(PLUGIN_DIR, "psim_host_demo.cpp"), "-o", demo + ".tmp", "-ldl"], verbose)
    os.replace(demo + ".tmp", demo)
    cli = os.path.join(PKG, "cli", "psim_sim")
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-I", INCLUDE,
          os.path.join(PKG, "cli", "psim_sim.cpp"), "-o", cli + ".tmp", "-L", PKG, "-lpsim_gpu",
          "-Wl,-rpath,$ORIGIN/.."], verbose)
    os.replace(cli + ".tmp", cli)


def build_oracle(verbose: bool = False) -> None:
    sys.path.insert(0, ROOT)
    from oracle.oracle import build_oracle as _b  # test infrastructure
    _b(force=True)
    if os.path.isdir(REFERENCE):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref"), f"REF={REFERENCE}",
              "-j8"], verbose)


def build_all(verbose: bool = False) -> None:
    build_product(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
