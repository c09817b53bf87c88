"""The learned (tau, delta) proposal on the device against the CPU oracle.

Bars (written per test): the device gradient within 2e-4 of the oracle's
float64 backprop (relative, normwise; the device accumulates in FP32); the
device forward within 2e-5 of the oracle's; the DM loop run with a network
follows exactly the schedule of (tau, delta) it records: the oracle run with
that schedule gives identical jpvt and ncols and R within 1e-12, and the
first step's recorded proposal equals the oracle network on the oracle's
inputs (within 1e-5); a trace run writes one record per DM step whose label
block is at least the defaults' block.
"""
import ctypes
import sys

import numpy as np
import pytest

import inputs
from helpers import assert_parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import qrdm_amd
    qrdm_amd.lib()
    return torch


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def vp(t):
    return ctypes.c_void_p(t.data_ptr())


def test_gradient_matches_oracle(gpu):
    import qrdm_amd
    from oracle import policy as pol
    torch = gpu
    rng = np.random.default_rng(5)
    for ns in (1, 37, 1000):
        th = pol.init_theta(ns).astype(np.float32)
        th[pol.B1:pol.W2] = rng.normal(0, 0.2, pol.NH)
        X = rng.random((ns, pol.NIN)).astype(np.float32)
        T = rng.random((ns, pol.NOUT)).astype(np.float32)
        d = [torch.from_numpy(v).cuda() for v in (th, X.ravel(), T.ravel())]
        g = torch.zeros(pol.NPARAM, dtype=torch.float32, device="cuda")
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        assert qrdm_amd.lib().qrdm_policy_grad(vp(d[0]), vp(d[1]), vp(d[2]), ns, vp(g), vp(loss),
                                               None) == 0
        torch.cuda.synchronize()
        lref, gref = pol.loss_grad(th.astype(np.float64), X, T)
        gd = g.cpu().numpy().astype(np.float64)
        err = np.linalg.norm(gd - gref) / np.linalg.norm(gref)
        say("ns %d: loss %.6f vs %.6f, gradient rel. error %.2e" % (ns, loss.item(), lref, err))
        assert err <= 2e-4
        assert abs(loss.item() - lref) <= 2e-5 * max(1.0, lref)
        for a, b in ((pol.W1, pol.B1), (pol.W2, pol.B2), (pol.W3, pol.B3), (pol.B3, pol.NPARAM)):
            assert np.linalg.norm(gd[a:b] - gref[a:b]) <= 2e-4 * np.linalg.norm(gref) + 1e-9


def test_predict_matches_oracle_forward(gpu):
    import qrdm_amd
    from qrdm_amd import learned
    from oracle import policy as pol
    torch = gpu
    rng = np.random.default_rng(8)
    th = pol.init_theta(4).astype(np.float32)
    net = learned.pack(torch, th)
    X = rng.random((300, pol.NIN)).astype(np.float32)
    xd = torch.from_numpy(X.ravel()).cuda()
    out = torch.zeros(600, dtype=torch.float32, device="cuda")
    assert qrdm_amd.lib().qrdm_policy_predict(vp(net), vp(xd), 300, vp(out), None) == 0
    torch.cuda.synchronize()
    y, _ = pol.forward(th.astype(np.float64), X)
    err = np.abs(out.cpu().numpy().reshape(300, 2) - y).max()
    say("forward max abs error %.2e" % err)
    assert err <= 2e-5


def test_adam_step_matches_numpy(gpu):
    import qrdm_amd
    from oracle import policy as pol
    torch = gpu
    rng = np.random.default_rng(9)
    th = pol.init_theta(7).astype(np.float32)
    X = rng.random((500, pol.NIN)).astype(np.float32)
    T = rng.random((500, pol.NOUT)).astype(np.float32)
    dth = torch.from_numpy(th.copy()).cuda()
    m1, m2 = torch.zeros_like(dth), torch.zeros_like(dth)
    xd, yd = torch.from_numpy(X.ravel()).cuda(), torch.from_numpy(T.ravel()).cuda()
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    ref = th.astype(np.float64)
    a = np.zeros_like(ref)
    b = np.zeros_like(ref)
    for t in range(1, 4):
        assert qrdm_amd.lib().qrdm_policy_adam(vp(dth), vp(m1), vp(m2), vp(xd), vp(yd), 500,
                                               1e-3, t, vp(loss), None) == 0
        _, g = pol.loss_grad(ref, X, T)
        a = 0.9 * a + 0.1 * g
        b = 0.999 * b + 0.001 * g * g
        ref = ref - 1e-3 * (a / (1 - 0.9 ** t)) / (np.sqrt(b / (1 - 0.999 ** t)) + 1e-8)
    torch.cuda.synchronize()
    diff = np.abs(dth.cpu().numpy() - ref)
    say("adam: max parameter difference after 3 steps %.2e" % diff.max())
    # steps of lr 1e-3 per parameter: a sign flip of a tiny gradient moves a
    # parameter by up to 2e-3 per step; almost all agree to float32 rounding
    assert np.quantile(diff, 0.99) <= 1e-5 and diff.max() <= 6e-3


def _random_net(torch, seed, scale=1.0):
    from qrdm_amd import learned
    from oracle import policy as pol
    th = pol.init_theta(seed).astype(np.float32) * scale
    return th, learned.pack(torch, th)


@pytest.mark.parametrize("kind", ["gauss", "graded"])
def test_dm_with_policy_follows_its_schedule(gpu, oracle, kind):
    import qrdm_amd
    from qrdm_amd import learned
    from oracle import policy as pol
    torch = gpu
    m, n = 1024, 512
    a = inputs.gaussian(m, n, 31) if kind == "gauss" else inputs.graded(m, n, 32)
    th, net = _random_net(torch, 12, 0.5)
    at = torch.from_numpy(np.ascontiguousarray(np.asfortranarray(a).T)).cuda()
    t = torch.zeros(n, dtype=torch.float64, device="cuda")
    jp = np.zeros(n, np.int32)
    nc = np.zeros(n, np.int32)
    info, steps = qrdm_amd.dgeqrdm_dev(at, m, t, jp, nc, np.array([0.9, 0.15, 0.0]), trace=True,
                                       policy=net)
    torch.cuda.synchronize()
    assert info == 0
    taus = np.array([s["tau"] for s in steps])
    deltas = np.array([s["delta"] for s in steps])
    say("%s: %d steps, tau in [%.3f, %.3f], delta in [%.3f, %.3f]"
        % (kind, len(steps), taus.min(), taus.max(), deltas.min(), deltas.max()))
    lo, hi = learned.RANGES[0], learned.RANGES[1]
    assert (taus >= lo - 1e-6).all() and (taus <= hi + 1e-6).all()
    # the first proposal is the oracle network on the oracle's inputs
    w0 = np.linalg.norm(a, axis=0)
    f0 = pol.features(w0, 0, m, n)
    y0, _ = pol.forward(th.astype(np.float64), f0[None, :])
    tau0 = learned.RANGES[0] + (learned.RANGES[1] - learned.RANGES[0]) * y0[0, 0]
    del0 = learned.RANGES[2] + (learned.RANGES[3] - learned.RANGES[2]) * y0[0, 1]
    assert abs(taus[0] - tau0) <= 1e-5 and abs(deltas[0] - del0) <= 1e-5
    got = dict(info=info, a=np.asfortranarray(at.cpu().numpy().T), jpvt=jp, tau=t.cpu().numpy(),
               ncols=nc[:len(steps)].copy())
    ref = oracle.dgeqrdm(a, sched=(taus, deltas))
    assert_parity(got, ref)


def test_batched_policy_matches_plain_policy(gpu):
    """The batch kernel's proposal (same arithmetic) picks the same pivots."""
    import qrdm_amd
    torch = gpu
    th, net = _random_net(torch, 13, 0.5)
    mats = [inputs.gaussian(300, 200, 40 + i) for i in range(4)]
    plan = qrdm_amd.BatchPlan([x.shape for x in mats], nb=64)
    lib = qrdm_amd.lib()
    assert lib.qrdm_batch_plan_set_policy(plan._p, vp(net)) == 0
    buf = torch.zeros(plan.a_size, dtype=torch.float64, device="cuda")
    for i, x in enumerate(mats):
        o = int(plan.a_off[i])
        buf[o:o + x.size] = torch.from_numpy(np.asfortranarray(x).ravel(order="F")).cuda()
    tau = torch.zeros(plan.tau_size, dtype=torch.float64, device="cuda")
    jpb = torch.zeros(plan.col_size, dtype=torch.int32, device="cuda")
    ncb = torch.zeros(plan.col_size, dtype=torch.int32, device="cuda")
    info = torch.zeros(4, dtype=torch.int32, device="cuda")
    plan.run(buf, tau, jpb, ncb, info, graph=False)
    torch.cuda.synchronize()
    jpb = jpb.cpu().numpy()
    for i, x in enumerate(mats):
        at = torch.from_numpy(np.ascontiguousarray(np.asfortranarray(x).T)).cuda()
        t = torch.zeros(200, dtype=torch.float64, device="cuda")
        jp = np.zeros(200, np.int32)
        nc = np.zeros(200, np.int32)
        assert qrdm_amd.dgeqrdm_dev(at, 300, t, jp, nc, np.array([0.9, 0.15, 0.0]),
                                    policy=net)[0] == 0
        c = int(plan.col_off[i])
        assert np.array_equal(jpb[c:c + 200], jp), i


def test_trace_records(gpu, oracle):
    import qrdm_amd
    from qrdm_amd import learned
    from oracle import policy as pol
    torch = gpu
    mats = [inputs.graded(400, 160 + 40 * i, 60 + i) for i in range(3)]
    # 20 orthonormal columns and 20 partners at cosine 0.93 with them (an
    # orthogonal part of 0.37): delta 0.9 rejects the partners, delta >= 0.95
    # keeps all 40 in the first block
    q = np.linalg.qr(np.random.default_rng(5).standard_normal((400, 40)))[0]
    mats.append(np.asfortranarray(np.hstack([q[:, :20], 0.99 * (0.93 * q[:, :20] + np.sqrt(
        1 - 0.93 ** 2) * q[:, 20:])])))
    plan = qrdm_amd.BatchPlan([x.shape for x in mats], nb=64)
    buf = torch.zeros(plan.a_size, dtype=torch.float64, device="cuda")
    for i, x in enumerate(mats):
        o = int(plan.a_off[i])
        buf[o:o + x.size] = torch.from_numpy(np.asfortranarray(x).ravel(order="F")).cuda()
    tau = torch.zeros(plan.tau_size, dtype=torch.float64, device="cuda")
    jp = torch.zeros(plan.col_size, dtype=torch.int32, device="cuda")
    nc = torch.zeros(plan.col_size, dtype=torch.int32, device="cuda")
    info = torch.zeros(len(mats), dtype=torch.int32, device="cuda")
    cap = 2000
    out = torch.zeros(cap * 40, dtype=torch.float32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    rg = np.array(learned.RANGES, np.float32)
    th = np.array([0.9, 0.15, 0.0])
    r = qrdm_amd.lib().qrdm_batch_trace(plan._p, vp(buf), vp(tau), vp(jp), vp(nc), vp(info),
                                        th.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        vp(out), cap, vp(cnt),
                                        rg.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), None)
    assert r == 0
    torch.cuda.synchronize()
    recs = out.cpu().numpy().reshape(cap, 40)[:int(cnt.item())]
    ncs = nc.cpu().numpy()
    steps = 0
    for i, x in enumerate(mats):
        ref = oracle.dgeqrdm(x)
        c = int(plan.col_off[i])
        assert np.array_equal(ncs[c:c + len(ref["ncols"])], ref["ncols"])
        steps += len(ref["ncols"])
        # the step-0 record of this matrix: inputs as the oracle computes them
        r0 = recs[(recs[:, 38] == 0) & (recs[:, 39] == x.shape[1])]
        assert len(r0) == 1
        f0 = pol.features(np.linalg.norm(x, axis=0), 0, *x.shape)
        assert np.abs(r0[0, :34] - f0).max() <= 1e-6
        # its blocks: the default's and the best grid point's
        d0, b0 = _label_blocks(x)
        assert (r0[0, 36], r0[0, 37]) == (d0, b0), (i, r0[0, 36:38], d0, b0)
        if i == 3:
            assert (d0, b0) == (20, 40)
    assert len(recs) == steps
    assert (recs[:, 34:36] >= -1e-6).all() and (recs[:, 34:36] <= 1 + 1e-6).all()
    assert (recs[:, 37] >= recs[:, 36]).all()
    say("trace: %d records, label block / default block = %.3f"
        % (len(recs), recs[:, 37].sum() / recs[:, 36].sum()))


LABEL_TAU = (0.25, 0.15, 0.10, 0.05, 0.01)
LABEL_DELTA = (0.80, 0.90, 0.95, 0.98, 0.99, 0.999)


def _label_blocks(x, tau_def=0.15, delta_def=0.9, kmax=64):
    """The label scoring of a first DM step in numpy: the top-kmax column
    norms (value descending, position ascending) are the candidates; a grid
    point (tau, delta) keeps those above tau * max, accepts greedily in that
    order each whose |cos| with every accepted one is below delta, and is
    scored by the accepted columns' Cholesky factor cut at the first
    diagonal below tau_def * r_00 (the panel's truncation test).  Returns
    (the defaults' block, the largest block over the grid)."""
    nrm = np.linalg.norm(x, axis=0)
    order = sorted(range(len(nrm)), key=lambda c: (-nrm[c], c))[:min(kmax, *x.shape)]
    cand = nrm[order]
    G = x[:, order].T @ x[:, order]

    def block(tau, delta):
        nc = int(np.sum(cand > tau * cand[0]))
        acc = [0]
        for i in range(1, nc):
            if all(abs(G[q, i] / cand[q] / cand[i]) < delta for q in acc):
                acc.append(i)
        L = np.zeros((len(acc), len(acc)))
        for k in range(len(acc)):
            d = G[acc[k], acc[k]] - L[k, :k] @ L[k, :k]
            L[k, k] = np.sqrt(max(d, 0.0))
            if k > 0 and not L[k, k] >= tau_def * L[0, 0]:
                return k
            for t in range(k + 1, len(acc)):
                L[t, k] = (G[acc[t], acc[k]] - L[t, :k] @ L[k, :k]) / L[k, k] if L[k, k] > 0 else 0.0
        return len(acc)

    d = block(tau_def, delta_def)
    return d, max([d] + [block(t, dl) for t in LABEL_TAU for dl in LABEL_DELTA])


@pytest.mark.slow
def test_trained_policy_against_defaults(gpu):
    """The committed network (trained on the device on 6000 seeded 2048 x
    1024 matrices of the three families, python -m qrdm_amd.learned train
    --matrices 6000)
    against the default parameters on one 8192 x 4096 matrix per family: no
    loss in backward error or in the rank-revealing ratio min_k |R_kk| /
    sigma_k (within a factor 2 of the defaults'), no more than one extra DM
    step.  The gain (steps, time) is printed; it is measured, not asserted."""
    from qrdm_amd import learned
    torch = gpu
    net = learned.load_policy(torch)
    res = learned.evaluate(torch, net, reps=2, log=say)
    for fam, r in res.items():
        d, l = r["default"], r["learned"]
        say("%s: steps %d -> %d, %.2f -> %.2f ms, min |R_kk|/sigma_k %.3g -> %.3g"
            % (fam, d["steps"], l["steps"], d["ms"], l["ms"], d["rr_min"], l["rr_min"]))
        assert d["backward"] <= 1e-13 and l["backward"] <= 1e-13
        assert l["rr_min"] >= 0.5 * d["rr_min"]
        assert l["steps"] <= d["steps"] + 1
