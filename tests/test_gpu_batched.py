"""Batched DM-QR (qrdm_batch_* / qrdm_dgeqrdm_batched): one persistent
launch factorising many uneven matrices.

Bar: each matrix to the plain path's parity bar against the CPU oracle
(jpvt, ncols and info identical; R and the reflectors within 1e-12, see
tests/helpers.py) on generic inputs (Gaussian; graded over 8 decades).  On
spectra log-uniform down to 1e-10 the trailing block near the end holds
entries whose relative rounding is eps * sigma_1 / sigma_k (about 1e-7 at
1e-9), so pivots there are rounding-determined: those matrices are held to
the bar on the leading pivots whose singular values exceed 1e-6 (jpvt
identical, leading R block within 1e-12).  Graph replays are bit-identical
to the eager launch; the 4096-matrix uneven batch is compared with a
per-matrix loop of the plain device path.
"""
import sys

import numpy as np
import pytest

import inputs
from helpers import assert_parity, rel

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


def pack(torch, plan, mats):
    buf = torch.zeros(plan.a_size, dtype=torch.float64, device="cuda")
    for i, a in enumerate(mats):
        m, n = a.shape
        o = int(plan.a_off[i])
        buf[o:o + m * n] = torch.from_numpy(np.asfortranarray(a).ravel(order="F")).cuda()
    return buf


def outputs(torch, plan):
    return (torch.zeros(max(plan.tau_size, 1), dtype=torch.float64, device="cuda"),
            torch.zeros(plan.col_size, dtype=torch.int32, device="cuda"),
            torch.zeros(plan.col_size, dtype=torch.int32, device="cuda"),
            torch.full((len(plan.shapes),), 99, dtype=torch.int32, device="cuda"))


def unpack(plan, i, buf, tau, jp, nc, info):
    m, n = plan.shapes[i]
    o, t, c = int(plan.a_off[i]), int(plan.tau_off[i]), int(plan.col_off[i])
    a = buf[o:o + m * n].reshape(n, m).T.copy(order="F")
    ncols = nc[c:c + n]
    k, s = 0, 0
    while k < n and s < min(m, n) and ncols[k] > 0:
        s += int(ncols[k])
        k += 1
    return dict(info=int(info[i]), a=a, jpvt=jp[c:c + n].copy(), tau=tau[t:t + min(m, n)].copy(),
                ncols=ncols[:k].copy())


def run_batch(gpu, plan, buf, thres=(0.9, 0.15, 0.0), stop=0, graph=False):
    torch = gpu
    tau, jp, nc, info = outputs(torch, plan)
    plan.run(buf, tau, jp, nc, info, thres=thres, stop=stop, graph=graph)
    torch.cuda.synchronize()
    return buf.cpu().numpy(), tau.cpu().numpy(), jp.cpu().numpy(), nc.cpu().numpy(), info.cpu().numpy()


def _mixed_small(seed, count):
    """Uneven small matrices; leads[i] = None (full parity) or the number of
    leading pivots held to the bar (singular values above 1e-6)."""
    rng = np.random.default_rng(seed)
    mats, leads = [], []
    for i in range(count):
        m = int(rng.integers(64, 301))
        n = int(rng.integers(32, min(m, 200) + 1))
        if i % 3 == 0:
            mats.append(inputs.gaussian(m, n, 1000 + i))
            leads.append(None)
        elif i % 3 == 1:
            mats.append(inputs.graded(m, n, 2000 + i))
            leads.append(None)
        else:
            k = min(m, n)
            u = np.linalg.qr(rng.standard_normal((m, k)))[0]
            v = np.linalg.qr(rng.standard_normal((n, k)))[0]
            s = np.sort(10.0 ** rng.uniform(-10, 0, k))[::-1]
            mats.append((u * s) @ v.T)
            leads.append(int(np.sum(s > 1e-6)))
    return mats, leads


def check(got, ref, lead):
    if lead is None:
        assert_parity(got, ref)
        return
    assert got["info"] == ref["info"]
    assert np.array_equal(got["jpvt"][:lead], ref["jpvt"][:lead])
    assert rel(np.triu(got["a"][:lead, :lead]), np.triu(ref["a"][:lead, :lead])) <= 1e-12


@pytest.mark.parametrize("waves,rows", [(4, 0), (4, 128), (4, 64), pytest.param(8, 0, marks=pytest.mark.skip(
    reason="8-wave shape switched off: runs of it on a 4096-matrix batch ended in "
           "host-heap aborts of unidentified cause"))])
def test_batched_vs_oracle(gpu, oracle, waves, rows):
    """rows > 0: panels of at least that many rows take the Gram-based form."""
    import qrdm_amd
    torch = gpu
    mats, leads = _mixed_small(7, 36)
    plan = qrdm_amd.BatchPlan([a.shape for a in mats], nb=64)
    plan.set_shape(waves)
    plan.set_panel_rows(rows)
    assert plan.query()["waves"] == waves and plan.query()["panel_rows"] == rows
    say("batch of %d, waves %d, panel rows %d" % (len(mats), waves, rows))
    buf, tau, jp, nc, info = run_batch(gpu, plan, pack(torch, plan, mats))
    for i, a in enumerate(mats):
        got = unpack(plan, i, buf, tau, jp, nc, info)
        ref = oracle.dgeqrdm(a)
        try:
            check(got, ref, leads[i])
        except AssertionError:
            say("matrix", i, a.shape, "lead", leads[i])
            raise


@pytest.mark.parametrize("rows", [0, 64])
def test_batched_nb_and_stop_modes(gpu, oracle, rows):
    """Rank-20 matrices with noise: the panels truncate (inside the
    Gram-based form too, rows = 64), and the stop modes end early."""
    import qrdm_amd
    torch = gpu
    rng = np.random.default_rng(3)
    mats = [inputs.lowrank_noise(int(m), int(n), 20, 50 + i)
            for i, (m, n) in enumerate(zip(rng.integers(80, 260, 10), rng.integers(40, 80, 10)))]
    for nb in (16, 64):
        for stop, eta in ((0, 0.0), (1, 0.0), (3, 1e-8)):
            plan = qrdm_amd.BatchPlan([a.shape for a in mats], nb=nb)
            plan.set_panel_rows(rows)
            buf, tau, jp, nc, info = run_batch(gpu, plan, pack(torch, plan, mats),
                                               thres=(0.9, 0.15, eta), stop=stop)
            for i, a in enumerate(mats):
                got = unpack(plan, i, buf, tau, jp, nc, info)
                ref = oracle.dgeqrdm(a, nb=nb, stop=stop, eta=eta)
                if stop:
                    # columns past the stop are left unfinished, as in the reference
                    k = int(np.sum(ref["ncols"]))
                    assert np.array_equal(got["ncols"], ref["ncols"])
                    assert np.array_equal(got["jpvt"], ref["jpvt"])
                    assert rel(np.triu(got["a"][:k]), np.triu(ref["a"][:k])) <= 1e-12
                else:
                    assert_parity(got, ref)
            plan.close()


@pytest.mark.parametrize("rows", [0, 128])
def test_batched_graph_replays_match_eager(gpu, rows):
    import qrdm_amd
    torch = gpu
    mats, _ = _mixed_small(11, 20)
    plan = qrdm_amd.BatchPlan([a.shape for a in mats], nb=64)
    plan.set_panel_rows(rows)
    src = pack(torch, plan, mats)
    work = src.clone()
    eager = run_batch(gpu, plan, work, graph=False)
    tau, jp, nc, info = outputs(torch, plan)
    for rep in range(3):
        work.copy_(src)
        tau.zero_()
        jp.zero_()
        torch.cuda.synchronize()
        plan.run(work, tau, jp, nc, info, graph=True)
        torch.cuda.synchronize()
        got = (work.cpu().numpy(), tau.cpu().numpy(), jp.cpu().numpy(), nc.cpu().numpy(),
               info.cpu().numpy())
        for x, y in zip(got, eager):
            assert np.array_equal(x, y), rep
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays


def test_batched_nonfinite_and_bad_items(gpu):
    import qrdm_amd
    torch = gpu
    mats, _ = _mixed_small(5, 6)
    mats[2] = mats[2].copy()
    mats[2][3, 1] = np.nan
    plan = qrdm_amd.BatchPlan([a.shape for a in mats], nb=64)
    src = pack(torch, plan, mats)
    buf, tau, jp, nc, info = run_batch(gpu, plan, src.clone())
    assert list(info) == [0, 0, -13, 0, 0, 0]
    m, n = mats[2].shape
    o = int(plan.a_off[2])
    assert np.array_equal(buf[o:o + m * n], src.cpu().numpy()[o:o + m * n], equal_nan=True)
    with pytest.raises(ValueError):
        qrdm_amd.BatchPlan([(1100, 1025)])
    with pytest.raises(ValueError):
        qrdm_amd.BatchPlan([(10, 5)], nb=65)


def test_batched_autotune_keeps_results(gpu, oracle):
    """Autotuning (nb among 16, 32, 64 for a plan made with nb <= 0, panel
    rows among 0, 128, 256) leaves the input untouched and the results equal
    to the oracle's at the nb it picked."""
    import qrdm_amd
    torch = gpu
    mats, leads = _mixed_small(21, 24)
    plan = qrdm_amd.BatchPlan([a.shape for a in mats], nb=0)
    src = pack(torch, plan, mats)
    before = src.cpu().numpy()
    q = plan.autotune(src)
    say("autotune:", q)
    assert q["nb"] in (16, 32, 64) and q["waves"] == 4 and q["tuned"] == 9
    assert q["panel_rows"] in (0, 128, 256)
    assert np.array_equal(src.cpu().numpy(), before)
    buf, tau, jp, nc, info = run_batch(gpu, plan, src, graph=True)
    for i, a in enumerate(mats):
        check(unpack(plan, i, buf, tau, jp, nc, info), oracle.dgeqrdm(a, nb=q["nb"]), leads[i])


def test_batched_gram_panel_tall_uneven(gpu, oracle):
    """Tall uneven matrices (m up to 1024, the batched workload's range) with
    the Gram-based panel on every panel of 128 rows or more, against the
    oracle: Gaussian and graded inputs at the full bar, low-rank ones
    (truncation inside the Gram-based form) on their resolved pivots."""
    import qrdm_amd
    torch = gpu
    rng = np.random.default_rng(17)
    mats, leads = [], []
    for i in range(24):
        m = int(rng.integers(300, 1025))
        n = int(rng.integers(64, min(m, 400) + 1))
        if i % 3 == 0:
            mats.append(inputs.gaussian(m, n, 3000 + i))
            leads.append(None)
        elif i % 3 == 1:
            mats.append(inputs.graded(m, n, 4000 + i))
            leads.append(None)
        else:
            r = int(rng.integers(10, 60))
            mats.append(inputs.lowrank_noise(m, n, r, 5000 + i, noise=0.0))
            leads.append(r)
    plan = qrdm_amd.BatchPlan([a.shape for a in mats], nb=64)
    plan.set_panel_rows(128)
    say("tall batch of %d" % len(mats))
    buf, tau, jp, nc, info = run_batch(gpu, plan, pack(torch, plan, mats))
    for i, a in enumerate(mats):
        got = unpack(plan, i, buf, tau, jp, nc, info)
        ref = oracle.dgeqrdm(a)
        try:
            check(got, ref, leads[i])
        except AssertionError:
            say("matrix", i, a.shape, "lead", leads[i])
            raise


def test_batched_eight_wave_shape_is_refused(gpu):
    import qrdm_amd
    plan = qrdm_amd.BatchPlan([(40, 20)], nb=64)
    with pytest.raises(ValueError, match="switched off"):
        plan.set_shape(8)
    assert plan.query()["waves"] == 4


@pytest.mark.slow
def test_config4_batch_vs_plain_loop(gpu, oracle):
    """4096 uneven matrices (m in [64, 1024], n in [32, min(m, 512)], spectra
    log-uniform over [1e-10, 1]) in one launch, against a loop of the plain
    device path over the same matrices.  Both paths are rounding-determined in
    the tails (see the module docstring), and both take the Gram-based panel
    on tall panels but sum its products in different orders, so the
    check is: every matrix factorised (info 0), the leading pivots
    (sigma > 1e-6) identical on all but at most 8 of 4096 matrices (measured:
    2), the leading R blocks of identical-pivot matrices within 1e-12; each
    matrix whose leading pivots differ is printed beside the oracle."""
    import bench
    import qrdm_amd
    torch = gpu
    shapes = bench.batch_shapes(4096, 4)
    plan = qrdm_amd.BatchPlan(shapes, nb=64)
    say("config4: generating 4096 matrices")
    src, spectra = bench.batch_fill(torch, plan, 4)
    torch.cuda.synchronize()
    say("config4: batched run")
    work = src.clone()
    tau, jp, nc, info = outputs(torch, plan)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    plan.run(work, tau, jp, nc, info, graph=True)
    ev[1].record()
    torch.cuda.synchronize()
    t_batch = ev[0].elapsed_time(ev[1])
    info_h = info.cpu().numpy()
    assert (info_h == 0).all()
    say("config4: per-matrix loop of the plain path")
    loop = src.clone()
    tau2 = torch.zeros_like(tau)
    jp2 = np.zeros(plan.col_size, np.int32)
    nc2 = np.zeros(plan.col_size, np.int32)
    torch.cuda.synchronize()
    ev[0].record()
    for i, (m, n) in enumerate(shapes):
        o, t, c = int(plan.a_off[i]), int(plan.tau_off[i]), int(plan.col_off[i])
        a = loop[o:o + m * n].view(n, m)
        jv = jp2[c:c + n]
        nv = nc2[c:c + n]
        r, _ = qrdm_amd.dgeqrdm_dev(a, m, tau2[t:t + min(m, n)], jv, nv,
                                    np.array([0.9, 0.15, 0.0]))
        assert r == 0
    ev[1].record()
    torch.cuda.synchronize()
    t_loop = ev[0].elapsed_time(ev[1])
    say("config4: batched %.1f ms, loop %.1f ms (%.1fx)" % (t_batch, t_loop, t_loop / t_batch))
    jp1 = jp.cpu().numpy()
    same = lead_same = 0
    diffR = []
    w1, w2 = work.cpu().numpy(), loop.cpu().numpy()
    for i, (m, n) in enumerate(shapes):
        c, o = int(plan.col_off[i]), int(plan.a_off[i])
        lead = int(np.sum(spectra[i] > 1e-6))
        same += int(np.array_equal(jp1[c:c + n], jp2[c:c + n]))
        ok = np.array_equal(jp1[c:c + lead], jp2[c:c + lead])
        lead_same += int(ok)
        if ok and i % 16 == 0:
            a1 = w1[o:o + m * n].reshape(n, m).T
            a2 = w2[o:o + m * n].reshape(n, m).T
            diffR.append(rel(np.triu(a1[:lead, :lead]), np.triu(a2[:lead, :lead])))
    say("config4: jpvt identical on %d of 4096 matrices, leading pivots (sigma > 1e-6) on %d; "
        "max leading-R difference (every 16th matrix) %.2e" % (same, lead_same, max(diffR)))
    for i, (m, n) in enumerate(shapes):
        c, o = int(plan.col_off[i]), int(plan.a_off[i])
        lead = int(np.sum(spectra[i] > 1e-6))
        if not np.array_equal(jp1[c:c + lead], jp2[c:c + lead]):
            a0 = src[o:o + m * n].cpu().numpy().reshape(n, m).T.copy(order="F")
            ref = oracle.dgeqrdm(a0)
            say("  matrix %d (%dx%d, lead %d): batch %s oracle, plain loop %s oracle"
                % (i, m, n, lead,
                   "matches" if np.array_equal(jp1[c:c + lead], ref["jpvt"][:lead]) else "differs from",
                   "matches" if np.array_equal(jp2[c:c + lead], ref["jpvt"][:lead]) else "differs from"))
    assert lead_same >= 4096 - 8
    assert max(diffR) < 1e-12


@pytest.mark.slow
def test_config4_graph_replay_matches_eager(gpu):
    """The 4096-matrix uneven batch: the hipGraph replay of the persistent
    launch gives bit-identical results to the eager launch, every matrix
    factorised (info 0)."""
    import bench
    import qrdm_amd
    torch = gpu
    shapes = bench.batch_shapes(4096, 4)
    plan = qrdm_amd.BatchPlan(shapes, nb=64)
    src, _ = bench.batch_fill(torch, plan, 4)
    res = []
    for graph in (False, True, True):
        work = src.clone()
        tau, jp, nc, info = outputs(torch, plan)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        plan.run(work, tau, jp, nc, info, graph=graph)
        e1.record()
        torch.cuda.synchronize()
        say("config4 %s launch: %.1f ms" % ("graph" if graph else "eager", e0.elapsed_time(e1)))
        assert int((info != 0).sum().item()) == 0
        res.append([x.cpu().numpy() for x in (work, tau, jp, nc)])
    for r in res[1:]:
        for x, y in zip(r, res[0]):
            assert np.array_equal(x, y)
