"""CPU restatement of the learned (tau, delta) proposal (test infrastructure
only: imported by tests/ alone, as the checker of the device kernels).

No counterpart in the reference: the network, its inputs and its training
loss are this project's own (qrdm_amd/csrc/learned.hip), restated here in
float64 numpy with the same bfloat16 rounding of the weights:
  inputs   32-bin histogram of r = work[c] / max (bin floor(32 r), clamped
           to 31, counts / (n - j)), j / min(m, n), min((m-j)/(n-j), 8) / 8;
  forward  h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = sigmoid(W3 h2 + b3)
           with bf16(W) in place of W;
  loss     (1/N) sum (y - t)^2, and its gradient by hand-written backprop.
Pinned by: the gradient against central finite differences of the loss
(tests/test_policy_oracle.py), the inputs against a direct count on small
vectors, the bfloat16 rounding against numpy's float32 bit patterns on
ties and carries.
"""
import numpy as np

NIN, NH, NOUT, NBINS = 34, 64, 2, 32
W1, B1 = 0, NH * NIN
W2 = B1 + NH
B2 = W2 + NH * NH
W3 = B2 + NH
B3 = W3 + NOUT * NH
NPARAM = B3 + NOUT


def bf16_round(x):
    """Nearest bfloat16 (ties to even) of float32 values, as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r.astype(np.uint32)).view(np.float32).astype(np.float64)


def bf16_bits(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def features(work, j, m, n):
    """The 34 inputs at step j from the partial norms work[j..n)."""
    w = np.asarray(work[j:n], np.float64)
    mx = w.max() if len(w) else 0.0
    r = w / mx if mx > 0 else np.zeros_like(w)
    b = np.clip((r * NBINS).astype(np.int64), 0, NBINS - 1)
    f = np.zeros(NIN)
    f[:NBINS] = np.bincount(b, minlength=NBINS)[:NBINS] / float(n - j)
    f[NBINS] = j / float(min(m, n))
    f[NBINS + 1] = min((m - j) / float(n - j), 8.0) / 8.0
    return f.astype(np.float32).astype(np.float64)


def unpack(theta):
    t = np.asarray(theta, np.float64)
    return (t[W1:B1].reshape(NH, NIN), t[B1:W2], t[W2:B2].reshape(NH, NH), t[B2:W3],
            t[W3:B3].reshape(NOUT, NH), t[B3:NPARAM])


def forward(theta, X, round_weights=True):
    """Sigmoid outputs (N x 2) and the activations, float64."""
    w1, b1, w2, b2, w3, b3 = unpack(theta)
    if round_weights:
        w1, w2, w3 = bf16_round(w1), bf16_round(w2), bf16_round(w3)
    X = np.asarray(X, np.float64)
    z1 = X @ w1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w2.T + b2
    h2 = np.maximum(z2, 0)
    z3 = h2 @ w3.T + b3
    y = 1.0 / (1.0 + np.exp(-z3))
    return y, (z1, h1, z2, h2, w1, w2, w3)


def loss_grad(theta, X, T, round_weights=True):
    """Loss and its gradient with respect to theta (the gradient flows through
    the rounded weights unchanged, as the device applies it to the masters)."""
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    n = X.shape[0]
    y, (z1, h1, z2, h2, w1, w2, w3) = forward(theta, X, round_weights)
    e = y - T
    loss = float(np.sum(e * e) / n)
    d3 = 2.0 * e / n * y * (1.0 - y)
    g = np.zeros(NPARAM)
    g[W3:B3] = (d3.T @ h2).ravel()
    g[B3:NPARAM] = d3.sum(0)
    dz2 = (d3 @ w3) * (z2 > 0)
    g[W2:B2] = (dz2.T @ h1).ravel()
    g[B2:W3] = dz2.sum(0)
    dz1 = (dz2 @ w2) * (z1 > 0)
    g[W1:B1] = (dz1.T @ X).ravel()
    g[B1:W2] = dz1.sum(0)
    return loss, g


def init_theta(seed):
    """He-style initialisation of the parameter vector."""
    rng = np.random.default_rng(seed)
    th = np.zeros(NPARAM, np.float32)
    th[W1:B1] = rng.normal(0, np.sqrt(2.0 / NIN), NH * NIN)
    th[W2:B2] = rng.normal(0, np.sqrt(2.0 / NH), NH * NH)
    th[W3:B3] = rng.normal(0, np.sqrt(1.0 / NH), NOUT * NH)
    return th
