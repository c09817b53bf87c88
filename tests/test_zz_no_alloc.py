"""The allocation-free device entry point (qrdm_dgeqrdm_dev_noalloc) through
the C ABI: its workspace is the library's fixed device array, and its
results must equal those of qrdm_dgeqrdm_dev, which allocates its own --
jpvt, ncols and info identical, R and the reflectors within 1e-12 -- on two
calls in a row, the second on a smaller matrix (the array keeps the first
call's contents).  A matrix whose workspace exceeds the array is refused
before any device work.
"""
import ctypes
import sys

import numpy as np
import pytest

import inputs
from helpers import rel

pytestmark = pytest.mark.gpu
ip = ctypes.POINTER(ctypes.c_int)
dp = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import qrdm_amd
    return torch, qrdm_amd.lib()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def call(torch, fn, a, lda_pad=0, delta=0.9, tau=0.15, nb=64):
    m, n = a.shape
    lda = m + lda_pad
    buf = np.zeros((n, lda))
    buf[:, :m] = a.T  # column-major m x n with leading dimension lda
    d_a = torch.from_numpy(buf).cuda()
    d_tau = torch.zeros(min(m, n), dtype=torch.float64, device="cuda")
    jp = np.zeros(n, np.int32)
    nc = np.zeros(n, np.int32)
    th = np.array([delta, tau, 0.0])
    info = fn(102, m, n, ctypes.c_void_p(d_a.data_ptr()), lda, jp.ctypes.data_as(ip),
              ctypes.c_void_p(d_tau.data_ptr()), nc.ctypes.data_as(ip), th.ctypes.data_as(dp),
              nb, None)
    torch.cuda.synchronize()
    out = d_a.cpu().numpy()
    return dict(info=info, a=np.asfortranarray(out[:, :m].T), pad=out[:, m:].copy(), jpvt=jp,
                tau=d_tau.cpu().numpy(), ncols=nc)


def same(got, ref):
    assert got["info"] == 0 and ref["info"] == 0, (got["info"], ref["info"])
    assert np.array_equal(got["jpvt"], ref["jpvt"])
    assert np.array_equal(got["ncols"], ref["ncols"])
    m, n = got["a"].shape
    k = min(m, n)
    R, Rr = np.triu(got["a"][:k]), np.triu(ref["a"][:k])
    assert rel(R, Rr) <= 1e-12
    assert rel(np.tril(got["a"], -1), np.tril(ref["a"], -1)) <= 1e-12
    assert np.abs(got["tau"] - ref["tau"]).max() <= 1e-12
    assert np.array_equal(got["pad"], ref["pad"])  # rows past m untouched


def test_noalloc_matches_dev_on_two_calls(gpu):
    torch, L = gpu
    # a tall Gram-panel case on the merged trailing pass, then a smaller one
    # with a padded leading dimension (lda % 4 != 0: the three-kernel update)
    cases = [(2048, 1536, 21, 0), (700, 500, 22, 3)]
    refs = []
    for m, n, seed, pad in cases:
        say("qrdm_dgeqrdm_dev", m, n, "seed", seed, "lda pad", pad)
        refs.append(call(torch, L.qrdm_dgeqrdm_dev, inputs.gaussian(m, n, seed), pad))
    for (m, n, seed, pad), ref in zip(cases, refs):
        say("qrdm_dgeqrdm_dev_noalloc", m, n, "seed", seed, "lda pad", pad)
        got = call(torch, L.qrdm_dgeqrdm_dev_noalloc, inputs.gaussian(m, n, seed), pad)
        same(got, ref)


def test_noalloc_refuses_oversized_workspace(gpu):
    torch, L = gpu
    # 8192 x 8192 needs about 100 MiB of workspace, above the 64 MiB array;
    # the call returns before touching the (small, unused) buffer
    d = torch.zeros(16, dtype=torch.float64, device="cuda")
    jp = np.zeros(8192, np.int32)
    nc = np.zeros(8192, np.int32)
    th = np.array([0.9, 0.15, 0.0])
    info = L.qrdm_dgeqrdm_dev_noalloc(102, 8192, 8192, ctypes.c_void_p(d.data_ptr()), 8192,
                                      jp.ctypes.data_as(ip), ctypes.c_void_p(d.data_ptr()),
                                      nc.ctypes.data_as(ip), th.ctypes.data_as(dp), 64, None)
    assert info == -14
    assert float(d.abs().sum()) == 0.0
