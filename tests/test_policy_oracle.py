"""CPU checks of the oracle's restatement of the learned (tau, delta)
proposal (oracle/policy.py): the hand-written gradient against central
finite differences, the bfloat16 rounding against exact bit patterns, the
inputs against a direct count.  No GPU."""
import numpy as np

from oracle import policy as pol


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    th = pol.init_theta(1).astype(np.float64) * 0.7
    th[pol.B1:pol.W2] = rng.normal(0, 0.3, pol.NH)  # biases away from the ReLU kinks
    th[pol.B2:pol.W3] = rng.normal(0, 0.3, pol.NH)
    X = rng.random((17, pol.NIN))
    T = rng.random((17, pol.NOUT))
    loss, g = pol.loss_grad(th, X, T, round_weights=False)
    idx = np.concatenate([rng.choice(np.arange(pol.W1, pol.B1), 25, replace=False),
                          np.arange(pol.B1, pol.B1 + 5),
                          rng.choice(np.arange(pol.W2, pol.B2), 25, replace=False),
                          np.arange(pol.B2, pol.B2 + 5), np.arange(pol.W3, pol.NPARAM)])
    h = 1e-6
    for i in idx:
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (pol.loss_grad(tp, X, T, False)[0] - pol.loss_grad(tm, X, T, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_rounded_gradient_is_the_gradient_at_the_rounded_weights():
    rng = np.random.default_rng(2)
    th = pol.init_theta(3).astype(np.float64)
    X, T = rng.random((9, pol.NIN)), rng.random((9, pol.NOUT))
    l1, g1 = pol.loss_grad(th, X, T, round_weights=True)
    rt = th.copy()
    for a, b in ((pol.W1, pol.B1), (pol.W2, pol.B2), (pol.W3, pol.B3)):
        rt[a:b] = pol.bf16_round(th[a:b])
    l2, g2 = pol.loss_grad(rt, X, T, round_weights=False)
    assert l1 == l2 and np.array_equal(g1, g2)


def test_bf16_rounding_known_answers():
    def f(bits):
        return np.array([bits], np.uint32).view(np.float32)
    cases = [
        (0x3F800000, 0x3F80),  # 1.0 exact
        (0x3F808000, 0x3F80),  # tie, even stays
        (0x3F818000, 0x3F82),  # tie, odd rounds up
        (0x3F808001, 0x3F81),  # just above the tie
        (0x3F7FFFFF, 0x3F80),  # carry into the exponent
        (0xBF818000, 0xBF82),  # negative tie
    ]
    for x, want in cases:
        assert int(pol.bf16_bits(f(x))[0]) == want, hex(x)
        assert pol.bf16_round(f(x))[0] == np.array([want << 16], np.uint32).view(np.float32)[0]


def test_features_direct_count():
    work = np.array([0.0, 9.0, 3.0, 0.2, 4.0, 4.0, 1.0])
    f = pol.features(work, 2, 10, 7)  # columns 2..6: 3, .2, 4, 4, 1, max 4
    want = np.zeros(34)
    for v in (3.0, 0.2, 4.0, 4.0, 1.0):
        want[min(31, int(v / 4.0 * 32))] += 1 / 5
    want[32] = 2 / 7
    want[33] = min(8 / 5, 8) / 8
    assert np.allclose(f, want.astype(np.float32), rtol=0, atol=1e-7)
