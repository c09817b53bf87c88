"""The verification helpers (qrdm_amd.auxil, counterparts of the reference's
auxil.py:6-200): closed forms on the CPU, and Q formation through the GPU
DORMQR on real DM factorisations."""
import numpy as np
import pytest

import inputs
from qrdm_amd import auxil


def test_rank_from_sv_closed_forms():
    # first value below tol at position 3 -> 4 (the reference's "+1")
    assert auxil.rank_from_sv(np.array([5.0, 1.0, 1e-3, 1e-12, 1e-20])) == 4
    # unordered values: the first position below tol counts
    assert auxil.rank_from_sv(np.array([1e-10, 3.0, 2.0])) == 1
    assert auxil.rank_from_sv(np.array([3.0, -1e-10, 2.0])) == 2
    # nothing below tol -> length
    assert auxil.rank_from_sv(np.array([3.0, 2.0, 1.0])) == 3
    assert auxil.rank_from_sv(np.array([3.0, 2.0, 1e-6]), tol=1e-5) == 3
    assert auxil.rank_from_sv(np.array([], dtype=float)) == 0


@pytest.mark.parametrize("m,n", [(7, 5), (5, 7), (6, 6)])
def test_getR_layouts(m, n):
    rng = np.random.default_rng(m * 10 + n)
    X = rng.standard_normal((m, n))           # the m x n matrix
    kmax = min(m, n)
    # 102: a C-ordered (n, m) buffer is the column-major m x n matrix
    buf = np.ascontiguousarray(X.T)
    R = auxil.getR(102, m, n, buf)
    if kmax == n:
        np.testing.assert_array_equal(R, np.triu(X))
    for k in range(1, kmax):
        R = auxil.getR(102, m, n, buf, k)
        exp = X.copy()
        exp[:, :k] = np.triu(X[:, :k])
        np.testing.assert_array_equal(R, exp)
    # 101: the buffer is X itself
    R = auxil.getR(101, m, n, X)
    if kmax == n:
        np.testing.assert_array_equal(R, np.triu(X))
    R = auxil.getR(101, m, n, X, 2)
    exp = X.copy()
    exp[:, :2] = np.triu(X[:, :2])
    np.testing.assert_array_equal(R, exp)
    # k = 0 and k > min(m, n) both mean min(m, n)
    np.testing.assert_array_equal(auxil.getR(101, m, n, X, 0), auxil.getR(101, m, n, X, 99))


def test_bad_layout_is_an_error():
    with pytest.raises(ValueError):
        auxil.getR(100, 3, 3, np.eye(3))


def _dm(gpu, X, nb=64):
    import QRDM
    m, n = X.shape
    A = np.ascontiguousarray(X.T).copy()
    jp = np.zeros(n, np.int32)
    tau = np.zeros(min(m, n))
    nc = np.zeros(n, np.int32)
    info = QRDM.QRDM(102, m, n, A, m, jp, tau, nc, np.array([0.9, 0.15]), nb)
    assert info == 0
    return A, jp, tau


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import qrdm_amd
    qrdm_amd.lib()
    return torch


@pytest.mark.gpu
@pytest.mark.parametrize("m,n", [(300, 200), (257, 257), (520, 130)])
def test_checkQR_on_dm_factorisation(gpu, m, n):
    X = inputs.gaussian(m, n, 7000 + m + n)
    A, jp, tau = _dm(gpu, X)
    Q, R, err, errQ = auxil.checkQR(102, m, n, A, np.ascontiguousarray(X.T), tau, jp)
    assert sorted(jp - 1) == list(range(n))
    nn = min(m, n)
    # checkQR applies Q to the first min(m, n) columns of I only (as the
    # reference's does), so Q is orthonormal there; errQ covers all of Q and
    # is small only when m <= n
    Qn = Q[:, :nn]
    assert np.linalg.norm(Qn.T @ Qn - np.eye(nn)) <= 1e-13 * nn
    if m <= n:
        assert errQ <= 1e-13 * m
    assert err <= 1e-13 * np.linalg.norm(X) * np.sqrt(n)
    # |R_00| is the norm of the first pivot column
    assert abs(R[0, 0]) == pytest.approx(np.linalg.norm(X[:, jp[0] - 1]), rel=1e-13)


@pytest.mark.gpu
def test_low_rank_recovers_a_rank_r_matrix(gpu):
    m, n, r = 400, 300, 40
    X = inputs.lowrank_noise(m, n, r, 911)  # rank 40 + 1e-10 noise
    A, jp, tau = _dm(gpu, X)
    R = auxil.getR(102, m, n, A)
    k = auxil.rank_from_sv(np.abs(np.diag(R)) / abs(R[0, 0]), tol=1e-9)
    # the first negligible |R_kk| is at position r: rank_from_sv gives r + 1
    assert k == r + 1
    Alow = auxil.low_rank(102, m, n, A, np.ascontiguousarray(X.T), tau, jp, r)
    Xp = X[:, jp - 1]
    assert np.linalg.norm(Alow - Xp) <= 1e-8 * np.linalg.norm(X)
