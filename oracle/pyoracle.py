"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It loads oracle/liboracle.so (the C restatement of the
reference's dgeqrdm, qrdm_oracle.c) and, when present, the reference compiled
from its own sources (oracle/_ref/libqrdm_ref.so).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF = os.path.join(HERE, "_ref", "libqrdm_ref.so")
REF_BLAS_TXT = os.path.join(HERE, "_ref", "blas_path.txt")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None
_ref_ok = None


class Step(ctypes.Structure):
    _fields_ = [("j", ctypes.c_int), ("kmax", ctypes.c_int), ("nc", ctypes.c_int),
                ("fjb", ctypes.c_int), ("fjb_cmp", ctypes.c_int), ("nswap", ctypes.c_int),
                ("maxnrm", ctypes.c_double), ("sel", ctypes.c_int * 64)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        _lib.oracle_dgeqrdm_ex.restype = ctypes.c_int
        _lib.oracle_dgeqrdm_ex.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                           ctypes.c_int, _ip, _dp, _ip, _dp, ctypes.c_int,
                                           ctypes.POINTER(Step), ctypes.c_int, _ip]
        _lib.oracle_dgeqrdm_sched.restype = ctypes.c_int
        _lib.oracle_dgeqrdm_sched.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                              ctypes.c_int, _ip, _dp, _ip, _dp, ctypes.c_int,
                                              _dp, _dp, ctypes.c_int, ctypes.POINTER(Step),
                                              ctypes.c_int, _ip]
        _lib.oracle_nrm2.restype = ctypes.c_double
        _lib.oracle_nrm2.argtypes = [ctypes.c_int, _dp]
        for name in ("oracle_ref_dgeqrdm",):
            getattr(_lib, name).restype = ctypes.c_int
            getattr(_lib, name).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                            ctypes.c_int, _ip, _dp, _ip, _dp, ctypes.c_int]
        _lib.oracle_ref_dgeqp3.restype = ctypes.c_int
        _lib.oracle_ref_dgeqp3.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                           ctypes.c_int, _ip, _dp]
        _lib.oracle_ref_open.restype = ctypes.c_int
        _lib.oracle_ref_open.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        _lib.oracle_ref_set_threads.argtypes = [ctypes.c_int]
        _lib.o_dlarfg.restype = ctypes.c_int
        _lib.o_dlarfg.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int]
        _lib.o_dlarf_left.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, _dp,
                                      ctypes.c_int, _dp]
        _lib.o_dlarft_fc.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp,
                                     ctypes.c_int]
        _lib.o_dlarfb_ltfc.restype = ctypes.c_int
        _lib.o_dlarfb_ltfc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                       ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int]
        _lib.o_norm_update.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                       ctypes.c_int, _dp, _dp, _dp, ctypes.c_double]
        _lib.o_permute_marked.restype = ctypes.c_int
        _lib.o_permute_marked.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip,
                                          _dp, _ip, _ip, ctypes.c_int, ctypes.c_int]
        _lib.o_dgeqr2_dm.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp,
                                     ctypes.c_double, _ip]
        _lib.oracle_form_q.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp,
                                       _dp, ctypes.c_int]
    return _lib


def ref_available():
    """True when the compiled reference can be opened (this container only)."""
    global _ref_ok
    if _ref_ok is None:
        _ref_ok = False
        if os.path.exists(REF) and os.path.exists(REF_BLAS_TXT):
            blas = open(REF_BLAS_TXT).read().strip()
            _ref_ok = lib().oracle_ref_open(REF.encode(), blas.encode()) == 0
    return _ref_ok


def ptr(x, t):
    return x.ctypes.data_as(t)


def _prep(A, jpvt, ncols_mode, n):
    a = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    jp = np.zeros(n, np.int32) if jpvt is None else np.array(jpvt, np.int32)
    nc = np.zeros(max(n, 1), np.int32)
    nc[0] = ncols_mode
    return a, jp, nc


def dgeqrdm(A, delta=0.9, tau=0.15, nb=64, stop=0, eta=0.0, jpvt=None, trace=False,
            impl="oracle", layout=102, sched=None):
    """Runs DM-QR on a copy of the column-major matrix A.

    impl="oracle" runs the C restatement, impl="ref" the compiled reference.
    Returns dict(info, a, jpvt, tau, ncols, steps); ncols is trimmed to the
    steps taken, jpvt is 1-based as the reference returns it.
    """
    m, n = A.shape
    a, jp, nc = _prep(A, jpvt, stop, n)
    t = np.zeros(min(m, n), np.float64)
    th = np.array([delta, tau, eta], np.float64)
    L = lib()
    steps = []
    if impl == "oracle":
        cap = n + 1
        tr = (Step * cap)()
        ns = ctypes.c_int(0)
        if sched is None:
            info = L.oracle_dgeqrdm_ex(layout, m, n, ptr(a, _dp), m, ptr(jp, _ip), ptr(t, _dp),
                                       ptr(nc, _ip), ptr(th, _dp), nb, tr, cap, ctypes.byref(ns))
        else:
            st_ = np.ascontiguousarray(sched[0], np.float64)
            sd_ = np.ascontiguousarray(sched[1], np.float64)
            info = L.oracle_dgeqrdm_sched(layout, m, n, ptr(a, _dp), m, ptr(jp, _ip),
                                          ptr(t, _dp), ptr(nc, _ip), ptr(th, _dp), nb,
                                          ptr(st_, _dp), ptr(sd_, _dp), len(st_), tr, cap,
                                          ctypes.byref(ns))
        nsteps = ns.value
        if trace:
            steps = [dict(j=s.j, kmax=s.kmax, nc=s.nc, fjb=s.fjb, fjb_cmp=s.fjb_cmp,
                          nswap=s.nswap, maxnrm=s.maxnrm, sel=list(s.sel[:min(s.fjb, 64)]))
                     for s in tr[:nsteps]]
    elif impl == "ref":
        if not ref_available():
            raise RuntimeError("compiled reference not available")
        nc_before = nc.copy()
        info = L.oracle_ref_dgeqrdm(layout, m, n, ptr(a, _dp), m, ptr(jp, _ip), ptr(t, _dp),
                                    ptr(nc, _ip), ptr(th, _dp), nb)
        nsteps = _count_steps(nc, nc_before, min(m, n), stop)
    else:
        raise ValueError(impl)
    return dict(info=info, a=a, jpvt=jp, tau=t, ncols=nc[:nsteps].copy(), steps=steps,
                nsteps=nsteps)


def _count_steps(nc, nc_before, minmn, stop):
    """Number of DM steps recorded in ncols (sum of widths reaches minmn, or
    the stop test ended the loop earlier)."""
    s, k = 0, 0
    while k < len(nc) and s < minmn and nc[k] > 0:
        s += int(nc[k])
        k += 1
    return k


def dgeqp3_ref(A):
    """LAPACKE_dgeqp3 of the BLAS the compiled reference links."""
    if not ref_available():
        raise RuntimeError("compiled reference not available")
    m, n = A.shape
    a = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    jp = np.zeros(n, np.int32)
    t = np.zeros(min(m, n))
    info = lib().oracle_ref_dgeqp3(102, m, n, ptr(a, _dp), m, ptr(jp, _ip), ptr(t, _dp))
    return dict(info=info, a=a, jpvt=jp, tau=t)


def form_q(a, tau, k=None):
    m, n = a.shape
    k = min(m, n) if k is None else k
    af = np.asfortranarray(a).copy(order="F")
    q = np.zeros((m, k), order="F")
    lib().oracle_form_q(m, k, ptr(af, _dp), m, ptr(tau, _dp), ptr(q, _dp), m)
    return q
