"""The library's host code under AddressSanitizer, on the CPU.

driver.hip is compiled with its host part instrumented (-Xarch_host
-fsanitize=address) and linked against tests/hostcheck/stub_runtime.cpp (the
HIP runtime calls on host memory, and kernel launchers that model the DM
step protocol the host reads back) instead of libamdhip64.
tests/hostcheck/replay.cpp then makes the C-ABI calls of the batched
measurement -- a 4096-matrix plan (shape requests, autotune, hipGraph
replays, an eager launch), one qrdm_dgeqrdm_dev call per matrix with host
jpvt/ncols arrays of exactly n entries, the plan's destruction -- and the
other host paths (fixed columns, wide steps, traced and option calls, the
allocation-free and host-pointer entry points).  Any host-side overrun or use
after free ends the run with an ASan report.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang++"


def batch_shapes(count, seed, mmin=64, mmax=1024, nmin=32, nmax=512):
    """The batched measurement's shapes: m uniform in [mmin, mmax], n uniform
    in [nmin, min(m, nmax)]."""
    rng = np.random.default_rng(seed)
    ms = rng.integers(mmin, mmax + 1, count)
    return [(int(m), int(rng.integers(nmin, min(int(m), nmax) + 1))) for m in ms]


def run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, "%s\n%s%s" % (" ".join(cmd), r.stdout[-3000:], r.stderr[-3000:])
    return r


@pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)), reason="no ROCm clang")
def test_host_paths_clean_under_asan(tmp_path):
    d = str(tmp_path)
    feat = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
    asan = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]
    run([HIPCC, "--offload-arch=gfx950"] + feat + ["-O1", "-g", "-fPIC"] + asan +
        ["-c", os.path.join(ROOT, "qrdm_amd", "csrc", "driver.hip"), "-o", "driver.o"], d)
    for src in ("stub_runtime.cpp", "replay.cpp"):
        run([HIPCC, "--cuda-host-only", "-x", "hip", "-O1", "-g"] + asan +
            ["-c", os.path.join(HERE, "hostcheck", src), "-o", src + ".o"], d)
    run([CLANG, "-fsanitize=address", "driver.o", "stub_runtime.cpp.o", "replay.cpp.o", "-o",
         "replay"], d)
    with open(os.path.join(d, "shapes.txt"), "w") as f:
        for m, n in batch_shapes(4096, 4):
            f.write("%d %d\n" % (m, n))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0")
    r = subprocess.run(["./replay", "shapes.txt"], cwd=d, capture_output=True, text=True,
                       timeout=600, env=env)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out, out[-4000:]
    assert r.returncode == 0 and "clean (0 failures)" in out, out[-4000:]
