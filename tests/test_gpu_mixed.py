"""Reduced-precision trailing updates (qrdm_dgeqrdm_dev_opt, trailing =
FP32 or split BF16x3) against the FP64 path and the compiled reference.

The forms change only the two products of C := C - V T^T V^T C.  "fp32" and
"bf16x3" are scheduled: FP64 until maxnrm sqrt(n - j) <= 1e-12 maxnrm0 /
(64 u_low), the lower precision after; "fp32_all" and "bf16x3_all" use the
lower precision in every step.  Measured:
  backward error  ||A P - Q R||_F / ||A||_F   (Q R formed by the GPU dormqr)
  orthogonality   ||Q^T Q - I||_F             (Q formed by the GPU dormqr)
  rank revealing  |R_kk| / sigma_k            (sigma known by construction)
Bars, written per test: the FP64 path and the scheduled forms backward
error <= 1e-12 (FP64 <= 1e-14 on the graded matrices); fp32_all <= 2e-6 and
bf16x3_all <= 1e-5 (the FP32 unit roundoff 6e-8 times a modest growth);
orthogonality <= 2 eps n for every form (the reflectors are computed in
FP64; the FP64 path measures about 0.2 eps n); |R_kk| / sigma_k of a
reduced form within a factor 2 of the FP64 path's (and of the compiled
reference's) for every sigma_k above 1e3 times that form's backward error.
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


def factor_and_measure(torch, A, s, trailing, steps_out=None):
    """DM-QR of A (m x n, row-major torch tensor) with the given trailing
    precision; returns (metrics, jpvt, ncols, seconds).  steps_out: a list
    that receives the per-step records (trace)."""
    import qrdm_amd
    m, n = A.shape
    at = A.T.contiguous()  # column-major m x n
    t = torch.zeros(min(m, n), dtype=torch.float64, device="cuda")
    jp = np.zeros(n, np.int32)
    nc = np.zeros(n, np.int32)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    info, steps = qrdm_amd.dgeqrdm_dev(at, m, t, jp, nc, np.array([0.9, 0.15, 0.0]),
                                       trailing=trailing, trace=steps_out is not None)
    ev1.record()
    torch.cuda.synchronize()
    assert info == 0, (info, qrdm_amd.last_error())
    if steps_out is not None:
        steps_out.extend(steps)
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


BARS = {"fp64": 1e-14, "fp32": 1e-12, "bf16x3": 1e-12, "fp32_all": 2e-6, "bf16x3_all": 1e-5}
EPS = 2.0 ** -52


def _check_forms(torch, A, s, label):
    res = {}
    for form in ("fp64", "fp32", "bf16x3", "fp32_all", "bf16x3_all"):
        met, jp, nc, sec = factor_and_measure(torch, A, s, form)
        res[form] = met
        k = int(np.count_nonzero(nc))
        say("%s %s: %.3f s, %d steps, backward %.3e, orth %.3e, |R_kk|/sigma_k in [%.3g, %.3g]"
            % (label, form, sec, k, met["backward"], met["orth"], met["rr"].min(), met["rr"].max()))
        assert met["backward"] <= BARS[form], (form, met["backward"])
        assert met["orth"] <= 2 * EPS * A.shape[0], (form, met["orth"])
    sv = s[:len(res["fp64"]["rr"])].cpu().numpy()
    for form in ("fp32", "bf16x3", "fp32_all", "bf16x3_all"):
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
        for form in ("fp32_all", "bf16x3_all"):
            met, jp, nc, sec = factor_and_measure(torch, A, s, form)
            say("  %s backward %.3e orth %.3e" % (form, met["backward"], met["orth"]))
            assert met["backward"] <= BARS[form] and met["orth"] <= 2 * EPS * m


def lowrank_noise(torch, m, n, r, seed, noise=1e-10):
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    g1 = torch.randn(m, r, generator=gen, device="cuda", dtype=torch.float64)
    g2 = torch.randn(r, n, generator=gen, device="cuda", dtype=torch.float64)
    return g1 @ g2 + noise * torch.randn(m, n, generator=gen, device="cuda",
                                         dtype=torch.float64)


def test_scheduled_forms_switch_after_the_rank_cliff(gpu):
    """Low rank plus 1e-10 noise (config 2's family at 2048 x 2048, rank
    512): the scheduled forms run FP64 through the numerical-rank part --
    the same steps and pivots as the FP64 path there -- and the lower
    precision after the cliff, and still meet 1e-12."""
    torch = gpu
    import qrdm_amd
    A = lowrank_noise(torch, 2048, 2048, 512, 7)
    s = torch.ones(2048, dtype=torch.float64, device="cuda")
    st64 = []
    met64, jp64, nc64, sec64 = factor_and_measure(torch, A, s, "fp64", st64)
    assert met64["backward"] <= 1e-14
    assert all(t["prec"] == 0 for t in st64)
    for form, code in (("fp32", qrdm_amd.TRAILING["fp32"]), ("bf16x3", qrdm_amd.TRAILING["bf16x3"])):
        st = []
        met, jp, nc, sec = factor_and_measure(torch, A, s, form, st)
        low = [i for i, t in enumerate(st) if t["prec"] != 0]
        say("lowrank-2048 %s: %.3f s (fp64 %.3f s), %d steps, %d in the lower precision from "
            "column %d; backward %.3e, orth %.3e"
            % (form, sec, sec64, len(st), len(low), st[low[0]]["j"] if low else -1,
               met["backward"], met["orth"]))
        assert low, "the schedule never switched"
        first = low[0]
        assert all(t["prec"] == code for t in st[first:])  # sticky once switched
        assert st[first]["j"] >= 512  # not before the numerical rank is reached
        # FP64 up to the switch: the same steps and pivots as the FP64 path
        assert [t["fjb_cmp"] for t in st[:first]] == [t["fjb_cmp"] for t in st64[:first]]
        jlead = st[first]["j"]
        assert np.array_equal(jp[:jlead], jp64[:jlead])
        assert met["backward"] <= 1e-12, (form, met["backward"])
        assert met["orth"] <= 2 * EPS * 2048


def test_forms_against_compiled_reference_graded(gpu):
    """Graded 1024 x 1024 (s_i = 10^(-8 i / n)), built on the host: every
    form's backward error, orthogonality and |R_kk| / sigma_k beside the
    compiled reference's FP64 factorisation of the same matrix."""
    torch = gpu
    from oracle import pyoracle as po
    if not po.ref_available():
        pytest.skip("compiled reference (oracle/_ref) not present")
    n = 1024
    rng = np.random.default_rng(31)
    u, _ = np.linalg.qr(rng.standard_normal((n, n)))
    v, _ = np.linalg.qr(rng.standard_normal((n, n)))
    sv = 10.0 ** (-8.0 * np.arange(n) / n)
    A = np.asfortranarray((u * sv) @ v.T)
    ref = po.dgeqrdm(A, impl="ref")
    assert ref["info"] == 0
    q = po.form_q(ref["a"], ref["tau"])
    R = np.triu(ref["a"])
    be_ref = np.linalg.norm(q @ R - A[:, ref["jpvt"] - 1]) / np.linalg.norm(A)
    orth_ref = np.linalg.norm(q.T @ q - np.eye(n))
    rr_ref = np.abs(np.diag(R)) / sv
    say("graded-1024 reference: backward %.3e, orth %.3e, |R_kk|/sigma_k in [%.3g, %.3g]"
        % (be_ref, orth_ref, rr_ref.min(), rr_ref.max()))
    At = torch.from_numpy(np.ascontiguousarray(A)).cuda()
    st = torch.from_numpy(sv).cuda()
    for form in ("fp64", "fp32", "bf16x3", "fp32_all", "bf16x3_all"):
        met, jp, nc, sec = factor_and_measure(torch, At, st, form)
        keep = sv > 1e3 * max(met["backward"], be_ref)
        ratio = met["rr"][keep] / rr_ref[keep]
        say("graded-1024 %s: backward %.3e, orth %.3e, |R_kk|/sigma_k / reference's in "
            "[%.3f, %.3f] over %d k" % (form, met["backward"], met["orth"], ratio.min(),
                                         ratio.max(), keep.sum()))
        assert met["backward"] <= BARS[form]
        assert met["orth"] <= 2 * EPS * n
        assert ratio.min() > 0.5 and ratio.max() < 2.0
        if form == "fp64":
            assert np.array_equal(jp, ref["jpvt"])


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
