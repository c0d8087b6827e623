"""Build the native parts of tmc_amd in place.

    python -m tmc_amd.build            # engine + workload generators
    python -m tmc_amd.build --oracle   # also the parity checkers in oracle/

libtmc_dmc.so is compiled for gfx950 with hipcc.  FMA contraction is off
(the cell path reproduces the reference's float/double rounding points) and
packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
ENGINE_SO = os.path.join(HERE, "libtmc_dmc.so")
WORKLOADS_SO = os.path.join(HERE, "libtmc_workloads.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TMC_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops' is not a recognized feature" not in l
                    and "(ignore this warning" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


ENGINE_SOURCES = ["dmc_engine.hip", "dmc_server.hip", "dmc_tune.hip"]


def build_engine():
    srcs = [os.path.join(CSRC, f) for f in ENGINE_SOURCES]
    tmp = ENGINE_SO + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-pthread", "-o", tmp, *srcs])
    os.replace(tmp, ENGINE_SO)
    return ENGINE_SO


def build_workloads():
    src = os.path.join(CSRC, "workloads.cpp")
    tmp = WORKLOADS_SO + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", "-o", tmp, src])
    os.replace(tmp, WORKLOADS_SO)
    return WORKLOADS_SO


def build_oracle(with_reference=None):
    """liboracle.so always; oracle/_ref's libraries where the reference
    sources exist (cpp_dmc for libdmcref.so, mc for libtmcref.so)."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir, "liboracle.so"])
    ref = os.environ.get("TMC_REFERENCE_DIR", "/root/reference/cpp_dmc")
    ref_mc = os.environ.get("TMC_REFERENCE_MC_DIR", "/root/reference/mc")
    if with_reference is False:
        return
    if with_reference or os.path.isdir(ref):
        _run(["make", "-s", "-C", odir, "_ref/libdmcref.so", f"REF={ref}"])
    if os.path.isdir(ref_mc):
        _run(["make", "-s", "-C", odir, "_ref/libtmcref.so", f"REF_MC={ref_mc}"])


def build_all(oracle=True):
    build_engine()
    build_workloads()
    if oracle:
        build_oracle()


if __name__ == "__main__":
    build_all(oracle="--oracle" in sys.argv or "--all" in sys.argv)
    print("built", ENGINE_SO, WORKLOADS_SO)
