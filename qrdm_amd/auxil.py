"""Verification helpers around a QR factorisation held in LAPACK form.

Counterparts of the reference's numpy harness (auxil.py:6-200) with the same
names, arguments and results; Q is applied by this build's QRDM.DORMQR (the
GPU dormqr).  Arrays follow the reference's convention: with
matrix_layout = 102 a C-ordered numpy array of shape (n, m) is the
column-major m x n matrix, so `A.T` is the matrix itself.
"""
import numpy as np


def _qrdm():
    """This build's QRDM extension (ImportError when it has not been built)."""
    import QRDM
    return QRDM


def _dims(m, n, k):
    """(mm, nn, k) as the reference computes them: nn = min(m, n) columns of
    R are triangular, k is clipped to min(mm, nn) (0 meaning the maximum)."""
    mm = m
    nn = n if m >= n else m
    min_mn = min(mm, nn)
    if k == 0 or k > min_mn:
        k = min_mn
    return mm, nn, k


def _r_factor(matrix_layout, n, A, k):
    """R laid out as the reference's checkQR/getR build it
    (auxil.py:70-86, :137-152)."""
    if matrix_layout == 101:
        if k == n:
            return np.triu(A)
        R = np.zeros_like(A)
        R[:, :k] = np.triu(A[:, :k])
        R[:, k:] = A[:, k:].copy()
        return R
    if matrix_layout == 102:
        if k == n:
            return np.triu(A.T)
        R = np.zeros_like(A.T)
        R[:, :k] = np.triu(A[:k, :].T)
        R[:, k:] = A[k:, :].T
        return R
    raise ValueError("matrix_layout must be 101 or 102, got %r" % (matrix_layout,))


def rank_from_sv(s, tol=5.e-9):
    """Numerical rank from singular values or |diag R| (auxil.py:6-17).

    Returns the position of the first value with |s| < tol, plus one, or
    len(s) when no value is below tol.  The values need not be sorted.
    """
    small = np.where(np.absolute(s) < tol)[0]
    if len(small) > 0:
        return int(np.amin(small)) + 1
    return len(s)


def checkQR(matrix_layout, m, n, A, M, tau, jpvt, k=0, check_err=True, verbose=False):
    """Q, R and the absolute errors of a factorisation A*P = Q*R
    (auxil.py:20-105).

    A, tau: the factorisation as dgeqrdm/dgeqrf leave it; M: the original
    matrix (same layout); jpvt: 1-based pivots.  Returns (Q, R, err, errQ)
    with err = ||M[:, jpvt-1] - Q R||_F and errQ = ||I - Q^T Q||_F (both 0
    when check_err is False).
    """
    mm, nn, k = _dims(m, n, k)
    Q = np.eye(m)
    pvt = np.asarray(jpvt) - 1
    info = _qrdm().DORMQR(matrix_layout, mm, nn, k, A, m, tau, Q, m)
    if info != 0:
        raise RuntimeError("QRDM.DORMQR returned %d" % info)
    R = _r_factor(matrix_layout, n, A, k)
    if matrix_layout == 102:
        M = M.T
        Q = Q.T
    err = 0
    errQ = 0
    if check_err:
        errQ = np.linalg.norm(np.eye(m) - Q.T @ Q)
        if verbose:
            print("||Q.T*Q - I|| = ", errQ)
        if k < min(mm, nn) and verbose:
            print("||R_22||_2 = ", np.linalg.norm(R[k:, k:]))
        err = np.linalg.norm(M[:, pvt] - Q @ R)
        if verbose:
            print("||A[:,p] - Q*R|| = ", err)
    return Q, R, err, errQ


def getR(matrix_layout, m, n, A, k=0):
    """The R factor [[R11, R12], [0, R22]] with R11 upper triangular of
    order k (auxil.py:107-154)."""
    _, _, k = _dims(m, n, k)
    return _r_factor(matrix_layout, n, A, k)


def low_rank(matrix_layout, m, n, M, A, tau, jpvt, k):
    """Rank-k approximation Q [[R11, R12], [0, 0]] of A*P from the
    factorisation M, tau (auxil.py:156-200).  Only matrix_layout = 102 is
    defined (the reference builds R for 102 only).  Returns the m x n
    approximation as a C-ordered (m, n) array (the transpose of the
    column-major buffer)."""
    if matrix_layout != 102:
        raise ValueError("low_rank is defined for matrix_layout 102 only")
    _, _, k = _dims(m, n, k)
    if k == n:
        R = np.tril(M)
    else:
        R = np.zeros_like(M)
        R[:k, :] = np.tril(M[:k, :])
        R[k:, :k] = M[k:, :k]
    Alow = np.ascontiguousarray(R.copy())
    info = _qrdm().DORMQR(matrix_layout, m, n, k, M, m, tau, Alow, m)
    if info != 0:
        raise RuntimeError("QRDM.DORMQR returned %d" % info)
    return Alow.T
