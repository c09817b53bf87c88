"""Pins the CPU oracle (oracle/qrdm_oracle.c) -- CPU only.

1. against the reference compiled from its own sources (oracle/_ref), when it
   is present, on seeded inputs of many kinds;
2. against the committed golden fixtures written by that reference;
3. piece by piece against mathematics and LAPACK (scipy.linalg.lapack):
   dnrm2, dlarfg, dlarf, dlarft, dlarfb, the norm downdate (forced
   cancellation), the DM selection (brute force), permute_marked invariants.
"""
import ctypes
import itertools
import os

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

import inputs
from helpers import assert_parity, rel

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
dp = ctypes.POINTER(ctypes.c_double)
ip = ctypes.POINTER(ctypes.c_int)


def P(x, t=dp):
    return x.ctypes.data_as(t)


# ---------------------------------------------------------------- golden

def test_config1_golden(oracle):
    g = np.load(os.path.join(GOLD, "config1_ref.npz"))
    a = inputs.gaussian(1024, 512, int(g["seed"]))
    o = oracle.dgeqrdm(a)
    assert o["info"] == 0
    assert np.array_equal(o["jpvt"], g["jpvt"])
    assert np.array_equal(o["ncols"], g["ncols"])
    assert np.abs(o["tau"] - g["tau"]).max() < 1e-12
    R = np.triu(o["a"][:512])
    assert rel(np.diag(R), g["diag_r"]) < 1e-12
    assert rel(R[0], g["r_row0"]) < 1e-12
    assert rel(np.linalg.norm(R, axis=0), g["r_colnorms"]) < 1e-12
    assert rel(np.linalg.norm(np.tril(o["a"], -1), axis=0), g["v_colnorms"]) < 1e-12


def _small_cases():
    g = np.load(os.path.join(GOLD, "small_ref.npz"))
    out = []
    for i in range(int(g["ncases"])):
        m, n, nb, stop, seed = (int(x) for x in g["case%d_params" % i])
        delta, tau = (float(x) for x in g["case%d_dt" % i])
        ref = dict(a=g["case%d_a" % i], jpvt=g["case%d_jpvt" % i], tau=g["case%d_tau" % i],
                   ncols=g["case%d_ncols" % i], info=int(g["case%d_info" % i]))
        out.append((str(g["case%d_kind" % i]), m, n, nb, delta, tau, stop, seed, ref))
    return out


@pytest.mark.parametrize("case", _small_cases(), ids=lambda c: "%s-%dx%d-nb%d" % c[:4])
def test_small_golden(oracle, case):
    kind, m, n, nb, delta, tau, stop, seed, ref = case
    a = inputs.small_case(kind, m, n, seed)
    o = oracle.dgeqrdm(a, delta=delta, tau=tau, nb=nb, stop=stop, eta=1e-3)
    assert_parity(o, ref)


# ------------------------------------------------- compiled reference sweep

KINDS = ["gauss", "int", "zerocols", "dupcols", "scaled"]


@pytest.mark.parametrize("seed", range(40))
def test_against_compiled_reference(oracle, seed):
    if not oracle.ref_available():
        pytest.skip("compiled reference not present (oracle/_ref)")
    rng = np.random.default_rng(1000 + seed)
    kind = KINDS[seed % len(KINDS)]
    m, n = int(rng.integers(1, 160)), int(rng.integers(1, 160))
    nb = int(rng.choice([1, 2, 5, 8, 16, 32, 64]))
    delta = float(rng.choice([0.0, 0.5, 0.9, 1.0]))
    tau = float(rng.choice([0.0, 0.15, 0.5, 1.0]))
    stop = int(rng.choice([0, 0, 1, 2]))
    a = inputs.small_case(kind, m, n, 5000 + seed)
    o = oracle.dgeqrdm(a, delta=delta, tau=tau, nb=nb, stop=stop)
    r = oracle.dgeqrdm(a, delta=delta, tau=tau, nb=nb, stop=stop, impl="ref")
    assert_parity(o, r)


# --------------------------------------------------------- piecewise pins

def test_nrm2_scaling(oracle):
    L = oracle.lib()
    for scale in (1.0, 1e-170, 1e170, 1e-300, 1e300):
        x = np.random.default_rng(1).standard_normal(257) * scale
        ref = scale * np.linalg.norm(x / scale)
        assert abs(L.oracle_nrm2(len(x), P(x)) - ref) <= 4e-16 * 257 ** 0.5 * ref
    z = np.zeros(5)
    assert L.oracle_nrm2(5, P(z)) == 0.0


def test_dlarfg_vs_lapack(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(2)
    for n in (2, 3, 17, 100):
        for alpha in (1.5, -0.7, 0.0, -0.0):
            x = rng.standard_normal(n - 1)
            beta_l, v_l, tau_l = lapack.dlarfg(n, alpha, x.copy())
            al = np.array([alpha])
            xx = x.copy()
            tt = np.zeros(1)
            assert L.o_dlarfg(n, P(al), P(xx), P(tt), 0.0, 1) == 0
            if alpha == 0.0 and np.signbit(alpha):
                # f2c d_sign (src/dlarfg.c:21-26) tests b >= 0, so -0.0 is
                # '+' and beta = -||x||; Fortran SIGN gives +||x|| here.
                assert al[0] == -np.linalg.norm(x) and tt[0] == 1.0
                continue
            assert abs(al[0] - beta_l) <= 1e-15 * abs(beta_l)
            assert abs(tt[0] - tau_l) <= 1e-15
            assert np.allclose(xx, v_l, rtol=1e-14, atol=0)
    # xnorm == 0: tau = 0 and alpha kept (H = I)
    al, xx, tt = np.array([3.0]), np.zeros(4), np.array([9.0])
    L.o_dlarfg(5, P(al), P(xx), P(tt), 0.0, 1)
    assert tt[0] == 0.0 and al[0] == 3.0
    # order 1: tau = 0
    al, tt = np.array([-2.0]), np.array([9.0])
    L.o_dlarfg(1, P(al), P(xx), P(tt), 1.0, 5)
    assert tt[0] == 0.0 and al[0] == -2.0
    # truncation: only for idx_house > 1 and ||x|| < thres; nothing touched
    al, xx, tt = np.array([1.0]), np.full(3, 0.1), np.array([7.0])
    assert L.o_dlarfg(4, P(al), P(xx), P(tt), 1.0, 2) == 1
    assert tt[0] == 7.0 and al[0] == 1.0 and np.all(xx == 0.1)
    assert L.o_dlarfg(4, P(al), P(xx), P(tt), 1.0, 1) == 0
    # tiny column: the safmin rescaling branch
    al, xx, tt = np.array([1e-300]), np.array([1e-300, -2e-300]), np.zeros(1)
    beta_l, v_l, tau_l = lapack.dlarfg(3, 1e-300, np.array([1e-300, -2e-300]))
    L.o_dlarfg(3, P(al), P(xx), P(tt), 0.0, 1)
    assert abs(al[0] - beta_l) <= 1e-14 * abs(beta_l) and abs(tt[0] - tau_l) < 1e-14
    assert np.allclose(xx, v_l, rtol=1e-13)


def test_dlarf_closed_form(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(3)
    m, n = 37, 11
    v = rng.standard_normal(m)
    v[0] = 1.0
    v[-3:] = 0.0  # trailing zeros are trimmed
    tau = 1.3
    C = np.asfortranarray(rng.standard_normal((m, n)))
    C[:, -2:] = 0.0  # trailing zero columns are trimmed
    ref = (np.eye(m) - tau * np.outer(v, v)) @ C
    w = np.zeros(n)
    L.o_dlarf_left(m, n, P(v), tau, P(C), m, P(w))
    assert rel(C, ref) < 1e-14


def _random_reflectors(m, k, rng):
    a = np.asfortranarray(rng.standard_normal((m, k)))
    qr, tau, work, info = lapack.dgeqrf(a.copy())
    return np.asfortranarray(qr), tau


def test_dlarft_is_product_of_reflectors(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(4)
    for m, k in ((50, 7), (64, 64), (9, 3)):
        V, tau = _random_reflectors(m, k, rng)
        tau[k // 2] = 0.0  # an identity reflector
        T = np.zeros((k, k), order="F")
        L.o_dlarft_fc(m, k, P(V), m, P(tau), P(T), k)
        Vu = np.tril(V, -1)[:, :k] + np.eye(m, k)
        H = np.eye(m)
        for i in range(k):
            H = H @ (np.eye(m) - tau[i] * np.outer(Vu[:, i], Vu[:, i]))
        assert rel(np.eye(m) - Vu @ T @ Vu.T, H) < 1e-14
        assert np.all(np.tril(T, -1) == 0)


def test_dlarfb_applies_ht(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(5)
    m, k, n = 60, 9, 13
    V, tau = _random_reflectors(m, k, rng)
    T = np.zeros((k, k), order="F")
    L.o_dlarft_fc(m, k, P(V), m, P(tau), P(T), k)
    C = np.asfortranarray(rng.standard_normal((m, n)))
    Vu = np.tril(V, -1)[:, :k] + np.eye(m, k)
    ref = (np.eye(m) - Vu @ T @ Vu.T).T @ C
    assert L.o_dlarfb_ltfc(m, n, k, P(V), m, P(T), k, P(C), m) == 0
    assert rel(C, ref) < 1e-14
    # NaN checks, in the reference's order: C (-13), T (-11), V (-9)
    C[3, 4] = np.nan
    assert L.o_dlarfb_ltfc(m, n, k, P(V), m, P(T), k, P(C), m) == -13
    C[3, 4] = 0.0
    T2 = T.copy(order="F")
    T2[0, 1] = np.nan
    assert L.o_dlarfb_ltfc(m, n, k, P(V), m, P(T2), k, P(C), m) == -11
    V2 = V.copy(order="F")
    V2[20, 2] = np.nan
    assert L.o_dlarfb_ltfc(m, n, k, P(V2), m, P(T), k, P(C), m) == -9
    V2 = V.copy(order="F")
    V2[0, 2] = np.nan  # strict upper part of V is not V
    assert L.o_dlarfb_ltfc(m, n, k, P(V2), m, P(T), k, P(C), m) == 0


def test_norm_downdate_and_recompute(oracle):
    """Closed form of the Drmac-Bujanovic update, and the recompute branch
    when the remaining part is tiny (forced cancellation)."""
    L = oracle.lib()
    rng = np.random.default_rng(6)
    m, fjb, n = 40, 5, 6
    a = np.asfortranarray(rng.standard_normal((fjb + m, n)))
    a[fjb:, 2] *= 1e-9  # column 2: almost all of its norm lies in the R12 rows
    a[:, 4] = 0.0
    full = np.linalg.norm(a, axis=0)
    work = full.copy()
    work[4] = 0.0
    workold = full.copy()
    workold[4] = 0.0
    mx = np.zeros(1)
    tol3z = np.sqrt(2.0 ** -53)
    L.o_norm_update(m, n, fjb, P(a), fjb + m, P(work), P(workold), P(mx), tol3z)
    below = np.linalg.norm(a[fjb:], axis=0)
    for c in (0, 1, 3, 5):  # downdated: ||below|| = sqrt(full^2 - ||R12||^2)
        assert abs(work[c] - below[c]) < 1e-13 * full[c]
        assert workold[c] == full[c]
    assert abs(work[2] - below[2]) < 1e-15 * below[2]  # recomputed
    assert workold[2] == work[2]
    assert work[4] == 0.0 and workold[4] == 0.0  # zero columns skipped
    assert mx[0] == max(work)


def _dm_select_brute(norms, G, tau, delta, kmax):
    order = sorted(range(len(norms)), key=lambda i: (-norms[i], i))
    nc = 0
    while nc < min(len(order), kmax) and norms[order[nc]] > tau * norms[order[0]]:
        nc += 1
    sel = [order[0]]
    if nc > 1:
        for i in order[1:nc]:
            cos = [abs(G[s, i]) / (norms[s] * norms[i]) for s in sel]
            if max(cos) < delta and len(sel) < kmax:
                sel.append(i)
    return sel, nc


def test_dm_selection_brute_force(oracle):
    """The oracle's whole first DM step (trace) against a brute-force
    selection on tiny inputs, including exact ties."""
    for seed in range(60):
        rng = np.random.default_rng(seed)
        m, n = int(rng.integers(2, 12)), int(rng.integers(2, 12))
        a = np.asfortranarray(rng.integers(-2, 3, (m, n)).astype(float))
        a[:, 0] += 1.0  # never an all-zero matrix
        # (delta = 1 with exactly parallel columns makes |cos| == delta a
        # rounding tie, so the brute force keeps delta away from 1)
        tau, delta = float(rng.choice([0.0, 0.3, 0.9])), float(rng.choice([0.2, 0.7]))
        nb = int(rng.integers(1, 6))
        o = oracle.dgeqrdm(a, delta=delta, tau=tau, nb=nb, trace=True)
        norms = np.sqrt(np.sum(a * a, axis=0))  # exact for small integers
        sel, nc = _dm_select_brute(list(norms), a.T @ a, tau, delta, min(nb, m, n))
        s0 = o["steps"][0]
        assert s0["nc"] == nc
        assert s0["sel"] == sel


def test_permute_marked_places_selection_first(oracle):
    """Exhaustive over small selections: the accepted columns end in the
    leading positions, the result is a permutation, leading accepted columns
    that need not move stay, and jpvt/norms move with the columns."""
    L = oracle.lib()
    for ncols in range(1, 8):
        for nsel in range(1, ncols + 1):
            for sel in itertools.permutations(range(ncols), nsel):
                m = 3
                a = np.asfortranarray(np.tile(np.arange(ncols, dtype=float), (m, 1)))
                jp = np.arange(1, ncols + 1, dtype=np.int32)
                nr = np.arange(ncols, dtype=float) + 0.5
                mark = np.zeros(ncols, np.int32)
                mark[list(sel)] = 1
                idx = np.array(sel, np.int32)
                L.o_permute_marked(m, ncols, P(a), m, P(jp, ip), P(nr), P(mark, ip), P(idx, ip), nsel, 0)
                cols = a[0].astype(int)
                assert sorted(cols) == list(range(ncols))
                assert set(cols[:nsel]) == set(sel)
                assert np.array_equal(jp - 1, cols) and np.array_equal(nr - 0.5, cols)
                assert np.all(mark[:nsel] == 1) and np.all(mark[nsel:] == 0)
                for c in sel:
                    if c < nsel and ncols - 1 not in sel:
                        assert cols[c] == c


def test_oracle_factorisation_identities(oracle):
    """A P = Q R and orthogonality for a generic input."""
    a = inputs.gaussian(120, 80, 7)
    o = oracle.dgeqrdm(a, nb=16)
    q = oracle.form_q(o["a"], o["tau"])
    R = np.triu(o["a"][:80])
    perm = o["jpvt"] - 1
    assert rel(q @ R, a[:, perm]) < 1e-14
    assert np.linalg.norm(q.T @ q - np.eye(80)) < 1e-13
