"""Seeded input generators shared by the fixture script, the tests and the
bench (numpy PCG64 streams: identical on every machine with this numpy).

Matrices are returned column-major (Fortran order), float64.
"""
import numpy as np

CONFIG1_SEED = 20261202
CONFIG2_SEED = 20261203


def gaussian(m, n, seed):
    """i.i.d. N(0,1) entries (config 1 shape: 1024 x 512)."""
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((m, n)))


def lowrank_noise(m, n, r, seed, noise=1e-10):
    """A = G1 G2 + noise * N, G1 m x r, G2 r x n, all N(0,1) (config 2:
    8192 x 8192, r = 2048)."""
    rng = np.random.default_rng(seed)
    g1 = rng.standard_normal((m, r))
    g2 = rng.standard_normal((r, n))
    a = np.asfortranarray(g1 @ g2)
    a += noise * np.asfortranarray(rng.standard_normal((m, n)))
    return a


def graded(m, n, seed, decades=8.0):
    """A = U diag(s) V^T with s_i = 10^(-decades*i/k), U, V from QR of seeded
    Gaussians (config 3 family)."""
    rng = np.random.default_rng(seed)
    k = min(m, n)
    u, _ = np.linalg.qr(rng.standard_normal((m, k)))
    v, _ = np.linalg.qr(rng.standard_normal((n, k)))
    s = 10.0 ** (-decades * np.arange(k) / k)
    return np.asfortranarray((u * s) @ v.T)


def small_case(kind, m, n, seed):
    """Small inputs of several kinds for the parity sweeps."""
    rng = np.random.default_rng(seed)
    if kind == "gauss":
        return np.asfortranarray(rng.standard_normal((m, n)))
    if kind == "int":  # many exactly tied norms
        return np.asfortranarray(rng.integers(-2, 3, (m, n)).astype(np.float64))
    if kind == "zerocols":
        a = rng.standard_normal((m, n))
        a[:, rng.integers(0, n, max(1, n // 5))] = 0.0
        return np.asfortranarray(a)
    if kind == "dupcols":  # near-repeated columns: cosines within 1e-7 of 1
        b = rng.standard_normal((m, max(1, n // 3)))
        a = b[:, rng.integers(0, b.shape[1], n)] + 1e-4 * rng.standard_normal((m, n))
        return np.asfortranarray(a)
    if kind == "scaled":  # column norms over 12 decades
        a = rng.standard_normal((m, n)) * (10.0 ** rng.uniform(-6, 6, n))
        return np.asfortranarray(a)
    if kind == "lowrank":  # rank min(m,n)//3 plus 1e-6 noise
        r = max(1, min(m, n) // 3)
        a = rng.standard_normal((m, r)) @ rng.standard_normal((r, n))
        return np.asfortranarray(a + 1e-6 * rng.standard_normal((m, n)))
    raise ValueError(kind)
