"""CPU-side checks of the product: the C-ABI library loads and exports every
entry point include/qrdm_amd.h declares, the Python module "QRDM" exposes
the reference's five functions, argument validation matches the reference,
and the host index logic of the permutation matches the oracle's restatement
of permute_marked.  Nothing here launches GPU work."""
import ctypes
import itertools
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ip = ctypes.POINTER(ctypes.c_int)
dp = ctypes.POINTER(ctypes.c_double)


def header_functions():
    txt = open(os.path.join(ROOT, "include", "qrdm_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", txt)))


def test_library_exports_header():
    import qrdm_amd
    L = qrdm_amd.lib()
    names = header_functions()
    assert "dgeqrdm" in names and "qrdm_dgeqrdm_dev" in names
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(qrdm_amd.EXPORTS)
    assert L.qrdm_version().startswith(b"qrdm_amd")


def test_python_module_surface():
    import QRDM
    for f in ("init", "QRF", "QP3", "QRDM", "DORMQR"):
        assert callable(getattr(QRDM, f))
    assert QRDM.init() is None


def test_python_module_refuses_read_only_outputs():
    """The arrays the library writes (A, jpvt, tau, ncols; C of DORMQR) must
    be writable: a read-only one raises TypeError before any device work;
    thres (read only by the library) may be read-only."""
    import QRDM
    a = np.zeros(20)
    jp = np.zeros(4, np.int32)
    t = np.zeros(4)
    nc = np.zeros(4, np.int32)
    th = np.array([0.9, 0.15, 0.0])
    for name in ("a", "jp", "t", "nc"):
        args = dict(a=a.copy(), jp=jp.copy(), t=t.copy(), nc=nc.copy())
        args[name].setflags(write=False)
        with pytest.raises(TypeError, match="writable"):
            QRDM.QRDM(102, 5, 4, args["a"], 5, args["jp"], args["t"], args["nc"], th, 8)
    c = np.zeros(20)
    c.setflags(write=False)
    with pytest.raises(TypeError, match="writable"):
        QRDM.DORMQR(102, 5, 4, 2, a, 5, t, c, 5)
    th_ro = th.copy()
    th_ro.setflags(write=False)
    with pytest.raises(TypeError, match="writable"):  # the read-only A is what is refused
        a_ro = a.copy()
        a_ro.setflags(write=False)
        QRDM.QRDM(102, 5, 4, a_ro, 5, jp, t, nc, th_ro, 8)


@pytest.mark.parametrize("bad", [
    dict(layout=100), dict(m=0), dict(n=-1), dict(lda=3), dict(delta=1.5),
    dict(delta=-0.1), dict(tau=2.0), dict(nb=0)])
def test_argument_errors_return_minus_one(bad):
    """Illegal arguments return -1 (src/dgeqrdm_work.c:559-581) before any
    device work, through both the C ABI and the Python module."""
    import QRDM
    import qrdm_amd
    p = dict(layout=102, m=5, n=4, lda=5, delta=0.9, tau=0.15, nb=8)
    p.update(bad)
    a = np.zeros(40)
    jp = np.zeros(8, np.int32)
    t = np.zeros(8)
    nc = np.zeros(8, np.int32)
    th = np.array([p["delta"], p["tau"], 0.0])
    L = qrdm_amd.lib()
    info = L.dgeqrdm(p["layout"], p["m"], p["n"], a.ctypes.data_as(dp), p["lda"],
                     jp.ctypes.data_as(ip), t.ctypes.data_as(dp), nc.ctypes.data_as(ip),
                     th.ctypes.data_as(dp), p["nb"])
    assert info == -1
    assert QRDM.QRDM(p["layout"], p["m"], p["n"], a, p["lda"], jp, t, nc, th, p["nb"]) == -1


def test_permute_walk_matches_oracle(oracle):
    """The library's host walk performs the same swaps as the oracle's
    restatement of permute_marked, exhaustively on small selections."""
    import qrdm_amd
    L = qrdm_amd.lib()
    LO = oracle.lib()
    for ncols in range(1, 8):
        for nsel in range(1, ncols + 1):
            for sel in itertools.permutations(range(ncols), nsel):
                s = np.array(sel, np.int32)
                sw = np.zeros(4 * ncols + 4, np.int32)
                k = L.qrdm_permute_walk(ncols, s.ctypes.data_as(ip), nsel, sw.ctypes.data_as(ip))
                perm = list(range(ncols))
                for i in range(k):
                    p, q = sw[2 * i], sw[2 * i + 1]
                    perm[p], perm[q] = perm[q], perm[p]
                a = np.asfortranarray(np.arange(ncols, dtype=float)[None, :])
                jp = np.arange(ncols, dtype=np.int32)
                nr = np.zeros(ncols)
                mark = np.zeros(ncols, np.int32)
                mark[list(sel)] = 1
                nsw = LO.o_permute_marked(1, ncols, a.ctypes.data_as(dp), 1, jp.ctypes.data_as(ip),
                                          nr.ctypes.data_as(dp), mark.ctypes.data_as(ip),
                                          s.ctypes.data_as(ip), nsel, 0)
                assert perm == list(a[0].astype(int)), (ncols, sel)
                assert nsw == k


def test_permute_walk_rejects_bad_input():
    import qrdm_amd
    L = qrdm_amd.lib()
    sw = np.zeros(16, np.int32)
    for sel, ncols in (([3], 3), ([0, 0], 4), ([-1], 2)):
        s = np.array(sel, np.int32)
        assert L.qrdm_permute_walk(ncols, s.ctypes.data_as(ip), len(sel), sw.ctypes.data_as(ip)) == -1


def test_qp3_forward_is_host_lapack():
    """QRDM.QP3 forwards to the host LAPACKE_dgeqp3, as the reference does."""
    import QRDM
    import scipy.linalg.lapack as lapack
    a = np.asfortranarray(np.random.default_rng(1).standard_normal((30, 20)))
    b = a.copy(order="F")
    jp = np.zeros(20, np.int32)
    t = np.zeros(20)
    assert QRDM.QP3(102, 30, 20, b.T, 30, jp, t) == 0  # raw pointer: b's memory
    qr, jpv, tau, work, info = lapack.dgeqp3(a.copy(order="F"))
    assert np.array_equal(jp, jpv)
    assert np.allclose(b, qr, rtol=0, atol=1e-13) and np.allclose(t, tau, atol=1e-15)


def _closed_form_arrangement(ncols, sel):
    """The placement permute_marked produces, in the closed form the device
    kernel k_select computes (kernels.hip): final content of each position
    and the number of swaps."""
    fjb, jt = len(sel), ncols - 1
    S = set(sel)
    free = [q for q in range(fjb) if q not in S]
    has_a = jt in S and jt > 0
    far_a = has_a and jt >= fjb
    content = list(range(ncols))
    k, nswap = (1 if far_a else 0), 0
    for s in sel:
        if has_a and s == jt:
            continue
        if s >= fjb:
            f = free[k]
            k += 1
            content[f], content[s] = s, f
            nswap += 1
    if far_a:
        f0 = free[0]
        if 0 in S:
            content[0], content[f0], content[jt] = jt, 0, f0
            nswap += 2
        else:
            content[0], content[jt] = jt, 0
            nswap += 1
    elif has_a:
        content[0], content[jt] = jt, 0
        nswap += 1
    return content, nswap


def _walk_arrangement(ncols, sel):
    import qrdm_amd
    sw = np.zeros(2 * ncols + 4, np.int32)
    s = np.array(sel, np.int32)
    k = qrdm_amd.lib().qrdm_permute_walk(ncols, s.ctypes.data_as(ip), len(sel),
                                         sw.ctypes.data_as(ip))
    content = list(range(ncols))
    for i in range(k):
        a, b = sw[2 * i], sw[2 * i + 1]
        content[a], content[b] = content[b], content[a]
    return content, k


def test_permute_closed_form_matches_walk():
    """The device's closed-form placement equals the swap walk: exhaustively
    for up to 6 columns, and on random selections of up to 64 of 300."""
    import itertools
    for ncols in range(1, 7):
        for fjb in range(1, ncols + 1):
            for sel in itertools.permutations(range(ncols), fjb):
                assert _closed_form_arrangement(ncols, list(sel)) == \
                    _walk_arrangement(ncols, list(sel)), (ncols, sel)
    rng = np.random.default_rng(11)
    for _ in range(3000):
        ncols = int(rng.integers(1, 300))
        fjb = int(rng.integers(1, min(64, ncols) + 1))
        sel = [int(x) for x in rng.choice(ncols, fjb, replace=False)]
        if rng.random() < 0.4 and ncols - 1 not in sel:
            sel[int(rng.integers(fjb))] = ncols - 1
        if rng.random() < 0.4 and 0 not in sel:
            sel[int(rng.integers(fjb))] = 0
        if len(set(sel)) != fjb:
            continue
        assert _closed_form_arrangement(ncols, sel) == _walk_arrangement(ncols, sel), \
            (ncols, sel)
