"""qrdm_amd -- QR with Deviation Maximization pivoting on AMD Instinct MI355X.

The computation lives in libqrdm_amd.so (HIP kernels for gfx950 and a C++
host driver behind a C ABI, include/qrdm_amd.h).  This package loads it and
offers thin helpers for callers holding device tensors.  The Python drop-in
for the reference's module is the extension "QRDM" built next to this
package (import QRDM).  Loading fails loudly if the library is missing: there
is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqrdm_amd.so")

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Opts(ctypes.Structure):
    """qrdm_opts of include/qrdm_amd.h."""
    _fields_ = [("trailing", ctypes.c_int), ("policy", ctypes.c_void_p)]


TRAILING = {"fp64": 0, "fp32": 1, "bf16x3": 2, "fp32_all": 3, "bf16x3_all": 4}


class BatchItem(ctypes.Structure):
    """qrdm_batch_item of include/qrdm_amd.h."""
    _fields_ = [("m", ctypes.c_int), ("n", ctypes.c_int), ("lda", ctypes.c_int),
                ("reserved", ctypes.c_int), ("a_off", ctypes.c_longlong),
                ("tau_off", ctypes.c_longlong), ("col_off", ctypes.c_longlong)]


class Step(ctypes.Structure):
    """qrdm_step of include/qrdm_amd.h."""
    _fields_ = [("j", ctypes.c_int), ("kmax", ctypes.c_int), ("nc", ctypes.c_int),
                ("fjb", ctypes.c_int), ("fjb_cmp", ctypes.c_int), ("nswap", ctypes.c_int),
                ("panel", ctypes.c_int), ("maxnrm", ctypes.c_double), ("sel", ctypes.c_int * 64),
                ("tau", ctypes.c_double), ("delta", ctypes.c_double), ("prec", ctypes.c_int)]


# Every entry point include/qrdm_amd.h declares.
EXPORTS = ("dgeqrdm", "qrdm_dgeqrdm_dev", "qrdm_dgeqrdm_dev_noalloc", "qrdm_dgeqrdm_dev_trace",
           "qrdm_dgeqrdm_dev_opt",
           "qrdm_dgeqrf_dev",
           "qrdm_dormqr_dev", "qrdm_dgeqrf", "qrdm_dormqr", "qrdm_permute_walk",
           "qrdm_profile_enable", "qrdm_profile_select", "qrdm_profile_reset",
           "qrdm_profile_read", "qrdm_batch_plan_create", "qrdm_batch_plan_destroy",
           "qrdm_batch_plan_query", "qrdm_batch_plan_set_shape",
           "qrdm_batch_plan_set_panel_rows", "qrdm_batch_autotune",
           "qrdm_batch_plan_set_policy", "qrdm_batch_trace", "qrdm_policy_sizes",
           "qrdm_policy_grad", "qrdm_policy_adam", "qrdm_policy_pack", "qrdm_policy_predict", "qrdm_dgeqrdm_batched",
           "qrdm_last_error", "qrdm_version")

PROFILE_CLASSES = ("larfb_w", "norm_update", "select", "panel", "permute", "init_norms",
                   "larfb_apply", "larfb_sum")


def lib():
    """The loaded libqrdm_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libqrdm_amd.so is not built; run python -m qrdm_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        V = ctypes.c_void_p
        L.dgeqrdm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip,
                              _dp, _ip, _dp, ctypes.c_int]
        L.qrdm_dgeqrdm_dev.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V, ctypes.c_int,
                                       _ip, V, _ip, _dp, ctypes.c_int, V]
        L.qrdm_dgeqrdm_dev_noalloc.argtypes = L.qrdm_dgeqrdm_dev.argtypes
        L.qrdm_dgeqrdm_dev_trace.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V,
                                             ctypes.c_int, _ip, V, _ip, _dp, ctypes.c_int, V,
                                             ctypes.POINTER(Step), ctypes.c_int, _ip]
        L.qrdm_dgeqrdm_dev_opt.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, V,
                                           ctypes.c_int, _ip, V, _ip, _dp, ctypes.c_int,
                                           ctypes.POINTER(Opts), V, ctypes.POINTER(Step),
                                           ctypes.c_int, _ip]
        L.qrdm_batch_plan_create.argtypes = [ctypes.c_int, ctypes.POINTER(BatchItem),
                                             ctypes.c_int]
        L.qrdm_batch_plan_create.restype = V
        L.qrdm_batch_plan_destroy.argtypes = [V]
        L.qrdm_batch_plan_query.argtypes = [V, _ip, ctypes.c_int]
        L.qrdm_batch_plan_set_shape.argtypes = [V, ctypes.c_int]
        L.qrdm_batch_plan_set_panel_rows.argtypes = [V, ctypes.c_int]
        L.qrdm_batch_plan_set_policy.argtypes = [V, V]
        _fp = ctypes.POINTER(ctypes.c_float)
        L.qrdm_batch_trace.argtypes = [V, V, V, V, V, V, _dp, V, ctypes.c_int, V, _fp, V]
        L.qrdm_policy_sizes.argtypes = [_ip, ctypes.c_int]
        L.qrdm_policy_grad.argtypes = [V, V, V, ctypes.c_int, V, V, V]
        L.qrdm_policy_adam.argtypes = [V, V, V, V, V, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                       V, V]
        L.qrdm_policy_pack.argtypes = [V, _fp, V, V]
        L.qrdm_policy_predict.argtypes = [V, V, ctypes.c_int, V, V]
        L.qrdm_batch_autotune.argtypes = [V, V, _dp, V]
        L.qrdm_dgeqrdm_batched.argtypes = [V, V, V, V, V, V, _dp, ctypes.c_int, ctypes.c_int, V]
        L.qrdm_dgeqrf_dev.argtypes = [ctypes.c_int, ctypes.c_int, V, ctypes.c_int, V, V]
        L.qrdm_dormqr_dev.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, V,
                                      ctypes.c_int, V, V, ctypes.c_int, V]
        L.qrdm_dgeqrf.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp]
        L.qrdm_dormqr.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                  ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.qrdm_permute_walk.argtypes = [ctypes.c_int, _ip, ctypes.c_int, _ip]
        L.qrdm_profile_enable.argtypes = [ctypes.c_int]
        L.qrdm_profile_select.argtypes = [ctypes.c_uint]
        L.qrdm_profile_read.argtypes = [_dp, ctypes.c_int]
        L.qrdm_profile_read.restype = ctypes.c_int
        L.qrdm_last_error.restype = ctypes.c_char_p
        L.qrdm_version.restype = ctypes.c_char_p
        _lib = L
    return _lib


def last_error():
    return lib().qrdm_last_error().decode()


def dgeqrdm_dev(a, m, tau, jpvt, ncols, thres, nb=64, stream=None, trace=False,
                layout=102, trailing="fp64", policy=None):
    """DM-QR in place of a device matrix.

    a: contiguous float64 device tensor of shape (n, lda) -- its memory is the
    column-major m x n matrix with leading dimension lda >= m; tau: float64
    device tensor with min(m, n) entries; jpvt, ncols: int32 numpy arrays;
    thres: float64 numpy array [delta, tau(, eta)].  stream: hipStream_t as
    an int (torch.cuda.current_stream().cuda_stream), None = default stream.
    trailing: precision of the trailing update's products, "fp64" (the
    plain path); "fp32" or "bf16x3" (FP64 until the remaining block's norm
    bound lets the lower precision stay within 1e-12 of A, then the lower
    one); "fp32_all" or "bf16x3_all" (the lower precision in every step)
    (qrdm_dgeqrdm_dev_opt).  policy: a
    device tensor holding a qrdm_policy (qrdm_amd.policy), or None.
    Returns (info, steps), steps = per-step records when trace is set.
    """
    import numpy as np
    L = lib()
    if a.dim() != 2 or not a.is_contiguous() or a.element_size() != 8:
        raise ValueError("a must be a contiguous float64 (n, lda) tensor")
    n, lda = int(a.shape[0]), int(a.shape[1])
    if not (1 <= m <= lda):
        raise ValueError("need 1 <= m <= lda")
    s = ctypes.c_void_p(stream if stream is not None else 0)
    thres = np.ascontiguousarray(thres, dtype=np.float64)
    args = (layout, m, n, ctypes.c_void_p(a.data_ptr()), lda, jpvt.ctypes.data_as(_ip),
            ctypes.c_void_p(tau.data_ptr()), ncols.ctypes.data_as(_ip),
            thres.ctypes.data_as(_dp), nb, s)
    if trailing not in TRAILING:
        raise ValueError("trailing must be one of %s" % ", ".join(TRAILING))
    if trace or trailing != "fp64" or policy is not None:
        cap = max(1, min(m, n)) if trace else 0
        tr = (Step * max(cap, 1))()
        ns = ctypes.c_int(0)
        opts = Opts(TRAILING[trailing], None if policy is None else policy.data_ptr())
        info = L.qrdm_dgeqrdm_dev_opt(*args[:-1], ctypes.byref(opts), s,
                                      tr if trace else None, cap, ctypes.byref(ns))
        if not trace:
            return info, None
        steps = [dict(j=t.j, kmax=t.kmax, nc=t.nc, fjb=t.fjb, fjb_cmp=t.fjb_cmp, nswap=t.nswap,
                      panel=t.panel, maxnrm=t.maxnrm, sel=list(t.sel[:min(t.fjb, 64)]),
                      tau=t.tau, delta=t.delta, prec=t.prec)
                 for t in tr[:ns.value]]
        return info, steps
    return L.qrdm_dgeqrdm_dev(*args), None


class BatchPlan:
    """A batch of independent DM-QR factorisations run by one persistent
    kernel launch (qrdm_batch_* of include/qrdm_amd.h).

    shapes: list of (m, n); the matrices are packed one after another,
    column-major with lda = m (a_off[i]), tau at tau_off[i], jpvt and ncols
    at col_off[i].  nb: DM block 1..64, or <= 0 to let autotune() pick it.
    """

    def __init__(self, shapes, nb=64):
        import numpy as np
        self.shapes = [(int(m), int(n)) for m, n in shapes]
        cnt = len(self.shapes)
        self.items = (BatchItem * cnt)()
        self.a_off = np.zeros(cnt, np.int64)
        self.tau_off = np.zeros(cnt, np.int64)
        self.col_off = np.zeros(cnt, np.int64)
        a = t = c = 0
        for i, (m, n) in enumerate(self.shapes):
            self.items[i] = BatchItem(m, n, m, 0, a, t, c)
            self.a_off[i], self.tau_off[i], self.col_off[i] = a, t, c
            a += m * n
            t += min(m, n)
            c += n
        self.a_size, self.tau_size, self.col_size = a, t, c
        self._p = lib().qrdm_batch_plan_create(cnt, self.items, nb)
        if not self._p:
            raise ValueError(last_error())

    def close(self):
        if getattr(self, "_p", None):
            lib().qrdm_batch_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def query(self):
        import numpy as np
        out = np.zeros(5, np.int32)
        lib().qrdm_batch_plan_query(self._p, out.ctypes.data_as(_ip), 5)
        return dict(waves=int(out[0]), grid=int(out[1]), nb=int(out[2]), tuned=int(out[3]),
                    panel_rows=int(out[4]))

    def set_panel_rows(self, rows):
        """Panels with at least rows rows (and 2+ columns) take the Gram-based
        form, the others the column-by-column one; rows <= 0: column by
        column only.  A new plan starts at 128."""
        if lib().qrdm_batch_plan_set_panel_rows(self._p, rows) != 0:
            raise ValueError(last_error())

    def set_shape(self, waves):
        """Launch shape by hand: 4 waves per workgroup (the one shape built;
        8 is switched off)."""
        if lib().qrdm_batch_plan_set_shape(self._p, waves) != 0:
            raise ValueError(last_error())

    def autotune(self, a, thres=(0.9, 0.15), stream=None):
        """Times the candidate launches (nb in {16, 32, 64} for a plan made
        with nb <= 0; panel rows in {0, 128, 256}) on a scratch copy of the
        batch tensor a and keeps the fastest."""
        import numpy as np
        th = np.ascontiguousarray(list(thres) + [0.0] * (3 - len(thres)), dtype=np.float64)
        info = lib().qrdm_batch_autotune(self._p, ctypes.c_void_p(a.data_ptr()),
                                         th.ctypes.data_as(_dp),
                                         ctypes.c_void_p(stream if stream else 0))
        if info != 0:
            raise RuntimeError("qrdm_batch_autotune: %d %s" % (info, last_error()))
        return self.query()

    def run(self, a, tau, jpvt, ncols, info, thres=(0.9, 0.15, 0.0), stop=0, graph=True,
            stream=None):
        """Factorises the batch in place (float64 a, tau; int32 jpvt, ncols,
        info device tensors); asynchronous on stream."""
        import numpy as np
        th = np.ascontiguousarray(list(thres) + [0.0] * (3 - len(thres)), dtype=np.float64)
        r = lib().qrdm_dgeqrdm_batched(self._p, ctypes.c_void_p(a.data_ptr()),
                                       ctypes.c_void_p(tau.data_ptr()),
                                       ctypes.c_void_p(jpvt.data_ptr()),
                                       ctypes.c_void_p(ncols.data_ptr()),
                                       ctypes.c_void_p(info.data_ptr()), th.ctypes.data_as(_dp),
                                       stop, 1 if graph else 0,
                                       ctypes.c_void_p(stream if stream else 0))
        if r != 0:
            raise RuntimeError("qrdm_dgeqrdm_batched: %d %s" % (r, last_error()))


def profile(enable=None, reset=False, classes=None):
    """Reads (and optionally resets / toggles) the library's kernel-time
    accounting: {class: (ms, launches, work)}.  classes: names from
    PROFILE_CLASSES to time while enabled (None: all)."""
    import numpy as np
    L = lib()
    if classes is not None:
        mask = 0
        for c in classes:
            mask |= 1 << PROFILE_CLASSES.index(c)
        L.qrdm_profile_select(ctypes.c_uint(mask))
    elif enable is not None:
        L.qrdm_profile_select(ctypes.c_uint(0xffffffff))
    if enable is not None:
        L.qrdm_profile_enable(1 if enable else 0)
    out = np.zeros(3 * len(PROFILE_CLASSES))
    L.qrdm_profile_read(out.ctypes.data_as(_dp), len(PROFILE_CLASSES))
    if reset:
        L.qrdm_profile_reset()
    return {c: tuple(out[3 * i:3 * i + 3]) for i, c in enumerate(PROFILE_CLASSES)}
