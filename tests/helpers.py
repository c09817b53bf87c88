"""Comparison helpers shared by the parity tests.

Parity bar (the reference's own results on the same inputs):
  * info, ncols and jpvt identical (jpvt and ncols: optionally only a
    leading part, for inputs whose trailing pivots and truncation points are
    decided by rounding noise);
  * R (upper trapezoid) within tol normwise, relative to ||R||;
  * each reflector k (V column and tau) within tol after weighting by
    |R_kk| / max|R_ii|: perturbation theory bounds reflector k's forward
    error by eps * ||A|| / |R_kk|, so the weighted error is the
    conditioning-free quantity; for well-conditioned inputs (config 1) the
    unweighted errors are checked as well (strict=True).
"""
import numpy as np


def split_rv(a):
    return np.triu(a), np.tril(a, -1)


def rel(x, y):
    ny = np.linalg.norm(y)
    return np.linalg.norm(x - y) / (ny if ny > 0 else 1.0)


def reflector_errors(got, ref, k):
    """Weighted per-reflector errors (V columns and tau) of the first k."""
    a_g, a_r = got["a"], ref["a"]
    d = np.abs(np.diag(a_r)[:k])
    scale = d.max() if k and d.max() > 0 else 1.0
    w = d / scale
    Vg = np.tril(a_g, -1)[:, :k]
    Vr = np.tril(a_r, -1)[:, :k]
    ev = np.linalg.norm(Vg - Vr, axis=0) / np.maximum(np.linalg.norm(Vr, axis=0), 1.0)
    et = np.abs(got["tau"][:k] - ref["tau"][:k])
    return ev * w, et * w, ev, et


def assert_parity(got, ref, tol=1e-12, lead=None, strict=False):
    """With lead, only the resolved part is compared: jpvt, R, V and tau of
    the first lead reflectors, and ncols of the steps that end within them
    (past it, pivots and truncation points are decided by rounding)."""
    assert got["info"] == ref["info"], (got["info"], ref["info"])
    if lead is None:
        assert np.array_equal(got["ncols"], ref["ncols"]), (got["ncols"][:20], ref["ncols"][:20])
    else:
        nl = int(np.sum(np.cumsum(ref["ncols"]) <= lead))
        assert np.array_equal(got["ncols"][:nl], ref["ncols"][:nl]), (got["ncols"][:nl],
                                                                     ref["ncols"][:nl])
    jg, jr = np.asarray(got["jpvt"]), np.asarray(ref["jpvt"])
    if lead is None:
        assert np.array_equal(jg, jr), np.nonzero(jg != jr)[0][:10]
        Rg, _ = split_rv(got["a"])
        Rr, _ = split_rv(ref["a"])
        assert rel(Rg, Rr) <= tol, rel(Rg, Rr)
        k = int(np.sum(ref["ncols"]))
    else:
        assert np.array_equal(jg[:lead], jr[:lead]), np.nonzero(jg[:lead] != jr[:lead])[0][:10]
        Rg = np.triu(got["a"][:lead, :lead])
        Rr = np.triu(ref["a"][:lead, :lead])
        assert rel(Rg, Rr) <= tol, rel(Rg, Rr)
        k = lead
    wv, wt, ev, et = reflector_errors(got, ref, k)
    if k:
        assert wv.max() <= tol, (wv.max(), int(wv.argmax()))
        assert wt.max() <= tol, (wt.max(), int(wt.argmax()))
        if strict:
            assert ev.max() <= tol and et.max() <= tol, (ev.max(), et.max())
