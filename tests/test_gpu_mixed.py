"""Reduced-precision trailing updates (qrdm_dgeqrdm_dev_opt, trailing =
FP32 or split BF16x3) against the FP64 path on graded spectra.

The FP64 path is the one pinned to the oracle (test_gpu_parity.py); these
forms change only the two products of C := C - V T^T V^T C, so the reported
figures are measured against it and against the known singular values:
  backward error  ||A P - Q R||_F / ||A||_F   (Q R formed by the GPU dormqr)
  orthogonality   ||Q^T Q - I||_F             (Q formed by the GPU dormqr)
  rank revealing  |R_kk| / sigma_k            (sigma known by construction)
Bars, written per test: FP64 backward error <= 1e-14; FP32 <= 2e-6 and
BF16x3 <= 1e-5 (the FP32 unit roundoff 6e-8 times a modest growth);
orthogonality <= 2 eps n for every form (the reflectors are computed in
FP64; the FP64 path measures about 0.2 eps n);
|R_kk| / sigma_k of a reduced form within a factor 2 of the FP64 path's for
every sigma_k above 1e3 times that form's backward error.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


def _dormqr(torch, trans, at, t, c):
    """c := Q c (trans 0) in place; at (n, m) holds the reflectors of an m x n
    factorisation, c an (ncols, m) tensor (column-major m x ncols)."""
    import qrdm_amd
    n, m = at.shape
    k = min(m, n)
    info = qrdm_amd.lib().qrdm_dormqr_dev(trans, m, c.shape[0], k, ctypes.c_void_p(at.data_ptr()),
                                          m, ctypes.c_void_p(t.data_ptr()),
                                          ctypes.c_void_p(c.data_ptr()), m, None)
    assert info == 0, info


def orthogonal(torch, n, gen):
    """Q of the unpivoted QR (GPU dgeqrf + dormqr) of an n x n Gaussian."""
    import qrdm_amd
    g = torch.randn(n, n, generator=gen, device="cuda", dtype=torch.float64)  # column-major G^T
    t = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert qrdm_amd.lib().qrdm_dgeqrf_dev(n, n, ctypes.c_void_p(g.data_ptr()), n,
                                          ctypes.c_void_p(t.data_ptr()), None) == 0
    q = torch.eye(n, dtype=torch.float64, device="cuda")
    _dormqr(torch, 0, g, t, q)
    return q.T  # (n, n) row-major Q


def graded(torch, n, seed):
    """A = U diag(s) V^T with s_i = 10^(-8 i / n), U and V orthogonal."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    u = orthogonal(torch, n, gen)
    v = orthogonal(torch, n, gen)
    s = 10.0 ** (-8.0 * torch.arange(n, dtype=torch.float64, device="cuda") / n)
    return (u * s) @ v.T, s


def factor_and_measure(torch, A, s, trailing):
    """DM-QR of A (m x n, row-major torch tensor) with the given trailing
    precision; returns (metrics, jpvt, ncols, seconds)."""
    import qrdm_amd
    m, n = A.shape
    at = A.T.contiguous()  # column-major m x n
    t = torch.zeros(min(m, n), dtype=torch.float64, device="cuda")
    jp = np.zeros(n, np.int32)
    nc = np.zeros(n, np.int32)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    info, _ = qrdm_amd.dgeqrdm_dev(at, m, t, jp, nc, np.array([0.9, 0.15, 0.0]),
                                   trailing=trailing)
    ev1.record()
    torch.cuda.synchronize()
    assert info == 0, (info, qrdm_amd.last_error())
    sec = ev0.elapsed_time(ev1) / 1e3
    R = torch.triu(at.T)  # (m, n)
    qr = R.T.contiguous()
    _dormqr(torch, 0, at, t, qr)
    perm = torch.from_numpy(jp.astype(np.int64) - 1).cuda()
    be = (torch.linalg.norm(qr.T - A[:, perm]) / torch.linalg.norm(A)).item()
    q = torch.eye(m, dtype=torch.float64, device="cuda")
    _dormqr(torch, 0, at, t, q)
    q = q.T.contiguous()
    orth = torch.linalg.norm(q.T @ q - torch.eye(m, dtype=torch.float64, device="cuda")).item()
    rr = (torch.abs(torch.diagonal(R)) / s[:min(m, n)]).cpu().numpy()
    return dict(backward=be, orth=orth, rr=rr), jp, nc, sec


BARS = {"fp64": 1e-14, "fp32": 2e-6, "bf16x3": 1e-5}
EPS = 2.0 ** -52


def _check_forms(torch, A, s, label):
    res = {}
    for form in ("fp64", "fp32", "bf16x3"):
        met, jp, nc, sec = factor_and_measure(torch, A, s, form)
        res[form] = met
        k = int(np.count_nonzero(nc))
        say("%s %s: %.3f s, %d steps, backward %.3e, orth %.3e, |R_kk|/sigma_k in [%.3g, %.3g]"
            % (label, form, sec, k, met["backward"], met["orth"], met["rr"].min(), met["rr"].max()))
        assert met["backward"] <= BARS[form], (form, met["backward"])
        assert met["orth"] <= 2 * EPS * A.shape[0], (form, met["orth"])
    sv = s[:len(res["fp64"]["rr"])].cpu().numpy()
    for form in ("fp32", "bf16x3"):
        keep = sv > 1e3 * res[form]["backward"]
        ratio = res[form]["rr"][keep] / res["fp64"]["rr"][keep]
        assert keep.sum() > 0
        assert ratio.min() > 0.5 and ratio.max() < 2.0, (form, ratio.min(), ratio.max())
    return res


def test_mixed_forms_graded_2048(gpu):
    torch = gpu
    say("graded 2048x2048 seed 3")
    A, s = graded(torch, 2048, 3)
    _check_forms(torch, A, s, "graded-2048")


def test_mixed_forms_tall_and_wide(gpu):
    """Tall and wide shapes (ragged tiles of both mixed kernels)."""
    torch = gpu
    gen = torch.Generator(device="cuda")
    gen.manual_seed(11)
    for m, n in ((1500, 700), (700, 1300), (333, 333)):
        k = min(m, n)
        say("gaussian-scaled %dx%d" % (m, n))
        u = orthogonal(torch, m, gen)[:, :k]
        v = orthogonal(torch, n, gen)[:, :k]
        s = 10.0 ** (-4.0 * torch.arange(k, dtype=torch.float64, device="cuda") / k)
        A = (u * s) @ v.T
        for form in ("fp32", "bf16x3"):
            met, jp, nc, sec = factor_and_measure(torch, A, s, form)
            say("  %s backward %.3e orth %.3e" % (form, met["backward"], met["orth"]))
            assert met["backward"] <= BARS[form] and met["orth"] <= 2 * EPS * m


def test_mixed_leaves_fp64_path_unchanged(gpu):
    """trailing = fp64 through the options entry point is the plain path."""
    torch = gpu
    import qrdm_amd
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    A = torch.randn(600, 400, generator=gen, device="cuda", dtype=torch.float64)
    outs = []
    for use_opt in (False, True):
        at = A.T.contiguous()
        t = torch.zeros(400, dtype=torch.float64, device="cuda")
        jp = np.zeros(400, np.int32)
        nc = np.zeros(400, np.int32)
        if use_opt:
            L = qrdm_amd.lib()
            th = np.array([0.9, 0.15, 0.0])
            opts = qrdm_amd.Opts(0, None)
            ns = ctypes.c_int(0)
            info = L.qrdm_dgeqrdm_dev_opt(102, 600, 400, ctypes.c_void_p(at.data_ptr()), 600,
                                          jp.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                          ctypes.c_void_p(t.data_ptr()),
                                          nc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                          th.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 64,
                                          ctypes.byref(opts), None, None, 0, ctypes.byref(ns))
        else:
            info, _ = qrdm_amd.dgeqrdm_dev(at, 600, t, jp, nc, np.array([0.9, 0.15, 0.0]))
        torch.cuda.synchronize()
        assert info == 0
        outs.append((at.cpu().numpy(), t.cpu().numpy(), jp.copy(), nc.copy()))
    assert np.array_equal(outs[0][2], outs[1][2]) and np.array_equal(outs[0][3], outs[1][3])
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_mixed_rejects_unknown_precision(gpu):
    torch = gpu
    import qrdm_amd
    L = qrdm_amd.lib()
    at = torch.zeros(8, 8, dtype=torch.float64, device="cuda")
    t = torch.zeros(8, dtype=torch.float64, device="cuda")
    jp = np.zeros(8, np.int32)
    nc = np.zeros(8, np.int32)
    th = np.array([0.9, 0.15, 0.0])
    opts = qrdm_amd.Opts(7, None)
    info = L.qrdm_dgeqrdm_dev_opt(102, 8, 8, ctypes.c_void_p(at.data_ptr()), 8,
                                  jp.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                  ctypes.c_void_p(t.data_ptr()),
                                  nc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                  th.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 64,
                                  ctypes.byref(opts), None, None, 0, None)
    assert info == -1


@pytest.mark.slow
def test_config3_graded_16384(gpu):
    """16384^2, s_i = 10^(-8 i / n): FP32 and BF16x3 trailing updates against
    the FP64 path (backward error, orthogonality, |R_kk| / sigma_k)."""
    torch = gpu
    say("config3 16384x16384 graded, seed 2024: generating")
    A, s = graded(torch, 16384, 2024)
    _check_forms(torch, A, s, "config3-16384")
