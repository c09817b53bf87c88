"""DM steps wider than 64 columns (64 < nb <= 256) through the C ABI,
against the compiled reference (oracle/_ref, which takes any nb) where it is
present and the oracle (the C restatement, any nb) otherwise.

Bar: as test_gpu_parity.py -- jpvt, ncols and info identical; R within
1e-12 normwise; reflectors (V, tau) within 1e-12 weighted by |R_kk|/|R_00|,
and unweighted on the generic Gaussian input (strict).
"""
import ctypes
import sys

import numpy as np
import pytest

import inputs
from helpers import assert_parity

pytestmark = pytest.mark.gpu
ip = ctypes.POINTER(ctypes.c_int)
dp = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import qrdm_amd
    qrdm_amd.lib()
    return torch


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_host(a, delta=0.9, tau=0.15, nb=64, stop=0, eta=0.0, jpvt=None):
    import qrdm_amd
    m, n = a.shape
    af = np.asfortranarray(a).copy(order="F")
    jp = np.zeros(n, np.int32) if jpvt is None else np.array(jpvt, np.int32)
    nc = np.zeros(max(n, 1), np.int32)
    nc[0] = stop
    t = np.zeros(min(m, n))
    th = np.array([delta, tau, eta])
    info = qrdm_amd.lib().dgeqrdm(102, m, n, af.ctypes.data_as(dp), m, jp.ctypes.data_as(ip),
                                  t.ctypes.data_as(dp), nc.ctypes.data_as(ip),
                                  th.ctypes.data_as(dp), nb)
    k = s = 0
    while k < len(nc) and s < min(m, n) and nc[k] > 0:
        s += int(nc[k])
        k += 1
    return dict(info=info, a=af, jpvt=jp, tau=t, ncols=nc[:k].copy())


def reference(oracle, a, **kw):
    if oracle.ref_available():
        return oracle.dgeqrdm(a, impl="ref", **kw), "reference"
    return oracle.dgeqrdm(a, **kw), "oracle"


@pytest.mark.parametrize("nb", [96, 128, 256])
def test_wide_config1_input(gpu, oracle, nb):
    """The config-1 input (1024 x 512 N(0,1), seeded) with nb > 64."""
    a = inputs.gaussian(1024, 512, inputs.CONFIG1_SEED)
    say("config1 1024x512 nb", nb)
    got = run_host(a, nb=nb)
    ref, who = reference(oracle, a, nb=nb)
    say("  %d steps (%s %d), widest %d" % (len(got["ncols"]), who, len(ref["ncols"]),
                                           int(got["ncols"].max())))
    assert got["ncols"].max() > 64  # the steps are wider than the 64-column path
    assert_parity(got, ref, strict=True)


@pytest.mark.parametrize("case", [
    ("lowrank", 800, 600, 128, 0),    # truncation inside wide panels
    ("tall", 1500, 300, 256, 0),
    ("wide", 300, 700, 96, 0),
    ("stop1", 900, 700, 128, 1),      # stop mode 1 (eta = eps n)
])
def test_wide_shapes(gpu, oracle, case):
    kind, m, n, nb, stop = case
    if kind == "lowrank":
        a = inputs.lowrank_noise(m, n, 150, 5)
    elif kind == "stop1":
        a = inputs.lowrank_noise(m, n, 120, 6, noise=0.0)
    else:
        a = inputs.gaussian(m, n, 9 + m)
    say("%s %dx%d nb %d stop %d" % (kind, m, n, nb, stop))
    got = run_host(a, nb=nb, stop=stop)
    ref, who = reference(oracle, a, nb=nb, stop=stop)
    say("  %d steps (%s %d)" % (len(got["ncols"]), who, len(ref["ncols"])))
    if kind in ("lowrank", "stop1"):
        # past the numerical rank the pivots are decided by rounding: the
        # resolved part (steps ending at or before the rank) is compared
        d = np.abs(np.diag(ref["a"]))
        starts = np.concatenate([[0], np.cumsum(ref["ncols"])[:-1]])
        deep = [s for s in starts if s < len(d) and d[s] < 1e-9 * d[0]]
        lead = int(deep[0]) if deep else None
        assert_parity(got, ref, lead=lead)
    else:
        assert_parity(got, ref, strict=True)


def test_wide_fixed_columns(gpu, oracle):
    """Fixed leading columns (jpvt != 0 on entry) before wide DM steps,
    against the oracle: the reference is undefined behaviour with fixed
    columns (its sort keys are set at absolute offsets and read at relative
    ones), so the corrected semantics of the oracle are the bar."""
    a = inputs.gaussian(600, 400, 41)
    jp0 = np.zeros(400, np.int32)
    jp0[[3, 10, 11, 200]] = 1
    got = run_host(a, nb=128, jpvt=jp0)
    ref = oracle.dgeqrdm(a, nb=128, jpvt=jp0)
    assert_parity(got, ref, strict=True)


def test_wide_ragged_and_small(gpu, oracle):
    """k_max = min(nb, m - j, n - j) below nb, and matrices narrower than nb."""
    for (m, n, nb, seed) in [(200, 90, 256, 1), (70, 150, 100, 2), (257, 257, 129, 3)]:
        a = inputs.gaussian(m, n, seed)
        got = run_host(a, nb=nb)
        ref, who = reference(oracle, a, nb=nb)
        say("%dx%d nb %d: %d steps" % (m, n, nb, len(got["ncols"])))
        assert_parity(got, ref, strict=True)


def test_nb_above_256_and_noalloc_refused(gpu):
    """nb > 256 returns -10; the allocation-free entry point refuses nb > 64."""
    torch = gpu
    import qrdm_amd
    a = inputs.gaussian(64, 32, 1)
    assert run_host(a, nb=257)["info"] == -10
    d = torch.zeros(64 * 32, dtype=torch.float64, device="cuda")
    t = torch.zeros(32, dtype=torch.float64, device="cuda")
    jp = np.zeros(32, np.int32)
    nc = np.zeros(32, np.int32)
    th = np.array([0.9, 0.15, 0.0])
    info = qrdm_amd.lib().qrdm_dgeqrdm_dev_noalloc(102, 64, 32, ctypes.c_void_p(d.data_ptr()), 64,
                                                   jp.ctypes.data_as(ip),
                                                   ctypes.c_void_p(t.data_ptr()),
                                                   nc.ctypes.data_as(ip), th.ctypes.data_as(dp),
                                                   96, None)
    assert info == -10
