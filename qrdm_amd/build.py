"""Builds the MI355X DM-QR library in-tree.

    python -m qrdm_amd.build        (or qrdm_amd.build.build() from Python)

Outputs (git-ignored, shipped with the tree):
  qrdm_amd/libqrdm_amd.so                 HIP kernels + host driver, C ABI
  QRDM.cpython-*.so (repository root)     the Python module "QRDM"
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are kept out of every device code object.  The
# matrix-core accumulators are allocated in ordinary VGPRs (gfx950 reads and
# writes them from either register file); with the AGPR form the compiler
# copied every accumulator between the two files on each pass of the
# trailing update's main loop.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-mllvm", "-amdgpu-mfma-vgpr-form=1"]
LIB = os.path.join(HERE, "libqrdm_amd.so")


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def ext_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(ROOT, "QRDM" + suffix)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    # kernels.hip and panel.hip compile together through device.hip
    for src in ("device.hip", "driver.hip"):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", LIB] + objs)
    inc = sysconfig.get_paths()["include"]
    _run(["gcc", "-O2", "-fPIC", "-shared", "-Wall", "-I" + inc,
          os.path.join(CSRC, "qrdm_pymod.c"), "-o", ext_path(),
          "-L" + HERE, "-lqrdm_amd", "-Wl,-rpath,$ORIGIN/qrdm_amd", "-ldl"])
    if verbose:
        print("built", LIB, "and", ext_path())
    return LIB


if __name__ == "__main__":
    build(verbose=True)
