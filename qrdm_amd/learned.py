"""The learned (tau, delta) proposal: training-data generation from the
method's own steps, training on the device, and the packed network.

The network and its kernels live in libqrdm_amd.so (csrc/learned.hip,
include/qrdm_amd.h: qrdm_policy_*, qrdm_batch_trace); this module drives
them with torch device tensors as plumbing:

  traces   the batched DM-QR run in trace mode over seeded matrices of three
           spectrum families writes one record per DM step: the 34 inputs
           and the label, the (tau, delta) of a grid whose greedy acceptance
           gives the largest block the panel's truncation test (default tau)
           keeps;
  train    minibatch Adam on the device (qrdm_policy_adam), FP32 master
           weights, bfloat16 forward;
  policy   the packed qrdm_policy the DM loop reads (qrdm_opts.policy).

Families (seeded): geometric s_i = 10^(-d i/(k-1)), d ~ U[4, 14]; step,
three clusters at 1, 1e-4, 1e-9 (sizes from a uniform split, values
jittered by 10%); Kahan-type, A = U K D with K = diag(sin(t)^i)(I - cos(t) N)
(N strictly upper triangular ones), t = 1.2, D = diag(1 + 0.01 N(0,1)).
U has orthonormal columns from two passes of Cholesky QR of a seeded
Gaussian (m >= 2n here, so the Gaussian is well conditioned), V from the
Householder QR of one (batched on the device).
"""
import ctypes
import os

import numpy as np

from . import BatchPlan, lib, last_error

RANGES = (0.01, 0.40, 0.75, 0.999)  # tau_lo, tau_hi, delta_lo, delta_hi
FAMILIES = ("geometric", "step", "kahan")
HERE = os.path.dirname(os.path.abspath(__file__))
THETA_PATH = os.path.join(HERE, "policy_theta.npy")


def sizes():
    out = np.zeros(3, np.int32)
    lib().qrdm_policy_sizes(out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 3)
    return dict(nparam=int(out[0]), net_bytes=int(out[1]), rec=int(out[2]))


def _orthonormal(torch, g):
    """Orthonormal columns of the batched (b, m, k) Gaussian g: Cholesky QR,
    twice."""
    q = g
    for _ in range(2):
        gram = torch.bmm(q.transpose(1, 2).contiguous(), q)
        r = torch.linalg.cholesky(gram, upper=True)
        q = torch.linalg.solve_triangular(r, q, upper=True, left=False)
    return q


def _spectrum(family, k, rng):
    if family == "geometric":
        d = rng.uniform(4.0, 14.0)
        return 10.0 ** (-d * np.arange(k) / max(k - 1, 1))
    if family == "step":
        cuts = np.sort(rng.integers(1, k, 2))
        s = np.empty(k)
        s[:cuts[0]] = 1.0
        s[cuts[0]:cuts[1]] = 1e-4
        s[cuts[1]:] = 1e-9
        s *= 1.0 + 0.1 * rng.uniform(-1, 1, k)
        return np.sort(s)[::-1]
    raise ValueError(family)


def _kahan(torch, n, theta=1.2):
    sn, cs = np.sin(theta), np.cos(theta)
    k = -cs * np.triu(np.ones((n, n)), 1) + np.eye(n)
    k = (sn ** np.arange(n))[:, None] * k
    return torch.from_numpy(k).cuda()


def family_batch(torch, family, count, m, n, seed):
    """(count, m, n) float64 device tensor of the family, and the spectra
    (None for Kahan-type, whose singular values are not set by
    construction)."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    rng = np.random.default_rng(seed + 7)
    k = min(m, n)
    u = _orthonormal(torch, torch.randn(count, m, k, generator=gen, device="cuda",
                                        dtype=torch.float64))
    if family == "kahan":
        kk = _kahan(torch, n)
        d = 1.0 + 0.01 * torch.randn(count, 1, n, generator=gen, device="cuda",
                                     dtype=torch.float64)
        a = torch.bmm(u, (kk[:k, :].unsqueeze(0) * d).contiguous())
        return a, None
    # V is square when m >= n: Householder QR (a square Gaussian can be too
    # ill-conditioned for Cholesky QR)
    v = torch.linalg.qr(torch.randn(count, n, k, generator=gen, device="cuda",
                                    dtype=torch.float64)).Q
    spectra = np.stack([_spectrum(family, k, rng) for _ in range(count)])
    s = torch.from_numpy(spectra).cuda().unsqueeze(1)
    a = torch.bmm((u * s).contiguous(), v.transpose(1, 2).contiguous())
    return a, spectra


def collect_traces(torch, count_per_family, m=2048, n=1024, seed=0, chunk=256,
                   thres=(0.9, 0.15), log=None):
    """Training records (N x 40 float32, host) from the method's steps on
    count_per_family matrices of each family."""
    recs = []
    rb = sizes()["rec"]
    ranges = np.array(RANGES, np.float32)
    th = np.array([thres[0], thres[1], 0.0])
    for fi, fam in enumerate(FAMILIES):
        done = 0
        while done < count_per_family:
            b = min(chunk, count_per_family - done)
            a, _ = family_batch(torch, fam, b, m, n, seed + 1000 * fi + done)
            plan = BatchPlan([(m, n)] * b, nb=64)
            buf = a.transpose(1, 2).contiguous().view(-1)  # column-major, packed
            del a
            tau = torch.zeros(plan.tau_size, dtype=torch.float64, device="cuda")
            jp = torch.zeros(plan.col_size, dtype=torch.int32, device="cuda")
            nc = torch.zeros(plan.col_size, dtype=torch.int32, device="cuda")
            info = torch.zeros(b, dtype=torch.int32, device="cuda")
            cap = b * n
            out = torch.zeros(cap * rb, dtype=torch.float32, device="cuda")
            cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
            r = lib().qrdm_batch_trace(plan._p, ctypes.c_void_p(buf.data_ptr()),
                                       ctypes.c_void_p(tau.data_ptr()),
                                       ctypes.c_void_p(jp.data_ptr()),
                                       ctypes.c_void_p(nc.data_ptr()),
                                       ctypes.c_void_p(info.data_ptr()),
                                       th.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       ctypes.c_void_p(out.data_ptr()), cap,
                                       ctypes.c_void_p(cnt.data_ptr()),
                                       ranges.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       None)
            if r != 0:
                raise RuntimeError("qrdm_batch_trace: %s" % last_error())
            torch.cuda.synchronize()
            k = min(int(cnt.item()), cap)
            recs.append(out[:k * rb].view(k, rb).cpu().numpy())
            plan.close()
            done += b
            if log:
                log("traces: %s %d/%d matrices, %d records" % (fam, done, count_per_family, k))
    return np.concatenate(recs)


def train(torch, X, Y, epochs=20, batch=4096, lr=2e-3, seed=0, holdout=0.05, log=None):
    """Minibatch Adam on the device (qrdm_policy_adam).  Returns the trained
    parameters (float32 numpy) and per-epoch (train loss, held-out loss)."""
    f32 = ctypes.c_void_p
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(X))
    nh = int(len(X) * holdout)
    Xd = torch.from_numpy(np.ascontiguousarray(X[order], np.float32)).cuda()
    Yd = torch.from_numpy(np.ascontiguousarray(Y[order], np.float32)).cuda()
    Xv, Yv, Xt, Yt = Xd[:nh], Yd[:nh], Xd[nh:], Yd[nh:]
    th = torch.from_numpy(init_theta(seed)).cuda()
    m1, m2 = torch.zeros_like(th), torch.zeros_like(th)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    grad = torch.zeros_like(th)
    n, t, hist = len(Xt), 0, []
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(n)).cuda()
        xs, ys = Xt[perm].contiguous(), Yt[perm].contiguous()
        tot, cnt = 0.0, 0
        for b0 in range(0, n - batch // 4, batch):
            nb = min(batch, n - b0)
            t += 1
            r = lib().qrdm_policy_adam(f32(th.data_ptr()), f32(m1.data_ptr()), f32(m2.data_ptr()),
                                       f32(xs[b0:].data_ptr()), f32(ys[b0:].data_ptr()), nb, lr, t,
                                       f32(loss.data_ptr()), None)
            if r != 0:
                raise RuntimeError("qrdm_policy_adam: %s" % last_error())
            tot += float(loss.item()) * nb
            cnt += nb
        vl = float("nan")
        if nh > 0:
            lib().qrdm_policy_grad(f32(th.data_ptr()), f32(Xv.data_ptr()), f32(Yv.data_ptr()), nh,
                                   f32(grad.data_ptr()), f32(loss.data_ptr()), None)
            vl = float(loss.item())
        hist.append((tot / max(cnt, 1), vl))
        if log:
            log("train: epoch %d loss %.5f held-out %.5f" % (ep, hist[-1][0], vl))
    return th.cpu().numpy(), hist


def init_theta(seed):
    """He-style initial parameters (float32 numpy)."""
    nh, nin = 64, 34
    rng = np.random.default_rng(seed)
    p = sizes()["nparam"]
    th = np.zeros(p, np.float32)
    w1, b1 = 0, nh * nin
    w2 = b1 + nh
    b2 = w2 + nh * nh
    w3 = b2 + nh
    b3 = w3 + 2 * nh
    th[w1:b1] = rng.normal(0, np.sqrt(2.0 / nin), nh * nin)
    th[w2:b2] = rng.normal(0, np.sqrt(2.0 / nh), nh * nh)
    th[w3:b3] = rng.normal(0, np.sqrt(1.0 / nh), 2 * nh)
    return th


def pack(torch, theta):
    """The qrdm_policy of theta, as a device byte tensor."""
    nb = sizes()["net_bytes"]
    net = torch.zeros(nb, dtype=torch.uint8, device="cuda")
    th = torch.from_numpy(np.ascontiguousarray(theta, np.float32)).cuda()
    rg = np.array(RANGES, np.float32)
    r = lib().qrdm_policy_pack(ctypes.c_void_p(th.data_ptr()),
                               rg.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                               ctypes.c_void_p(net.data_ptr()), None)
    if r != 0:
        raise RuntimeError("qrdm_policy_pack: %s" % last_error())
    torch.cuda.synchronize()
    return net


def load_policy(torch, path=THETA_PATH):
    """The committed trained network, packed on the device."""
    return pack(torch, np.load(path, allow_pickle=False))


def label_stats(recs, families=FAMILIES):
    """What the label grid can gain over the defaults under the panel's
    truncation judge, from trace records (in family order, as
    collect_traces writes them): the fraction of steps where some grid
    point keeps a larger block, the mean blocks, and which of tau and delta
    the winning labels move."""
    tau = RANGES[0] + (RANGES[1] - RANGES[0]) * recs[:, 34].astype(np.float64)
    delta = RANGES[2] + (RANGES[3] - RANGES[2]) * recs[:, 35].astype(np.float64)
    d_blk, b_blk = recs[:, 36], recs[:, 37]
    gain = b_blk > d_blk
    out = dict(steps=int(len(recs)), frac_steps_improvable=float(gain.mean()),
               mean_block_default=float(d_blk.mean()), mean_block_best=float(b_blk.mean()),
               mean_gain_where_improvable=float((b_blk - d_blk)[gain].mean()) if gain.any() else 0.0,
               winners_tau_below_default=int(np.sum(gain & (tau < 0.15 - 1e-6))),
               winners_tau_above_default=int(np.sum(gain & (tau > 0.15 + 1e-6))),
               winners_delta_above_default=int(np.sum(gain & (delta > 0.9 + 1e-6))),
               winners_delta_below_default=int(np.sum(gain & (delta < 0.9 - 1e-6))))
    vals, cnt = np.unique(np.round(delta[gain], 3), return_counts=True)
    out["winning_delta"] = {str(v): int(c) for v, c in zip(vals, cnt)}
    return out


def evaluate(torch, policy_net, m=8192, n=4096, seed=777, reps=3, thres=(0.9, 0.15), log=None):
    """Default parameters against the learned proposal on one matrix of each
    family: DM steps, time per factorisation, backward error and the
    rank-revealing ratio |R_kk| / sigma_k (sigma by construction, or
    computed for the Kahan-type matrix)."""
    import qrdm_amd
    out = {}
    for fi, fam in enumerate(FAMILIES):
        a, spectra = family_batch(torch, fam, 1, m, n, seed + fi)
        a = a[0]
        if spectra is None:
            sv = torch.linalg.svdvals(a).cpu().numpy()
        else:
            sv = spectra[0]
        res = {}
        for label, net in (("default", None), ("learned", policy_net)):
            ts = []
            for _ in range(reps):
                at = a.T.contiguous()
                t = torch.zeros(min(m, n), dtype=torch.float64, device="cuda")
                jp = np.zeros(n, np.int32)
                nc = np.zeros(n, np.int32)
                torch.cuda.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                info, _ = qrdm_amd.dgeqrdm_dev(at, m, t, jp, nc, np.array([thres[0], thres[1], 0.0]),
                                               policy=net)
                e1.record()
                torch.cuda.synchronize()
                if info != 0:
                    raise RuntimeError("dgeqrdm_dev: %d" % info)
                ts.append(e0.elapsed_time(e1))
            steps = int(np.count_nonzero(nc))
            R = torch.triu(at.T)
            qr = R.T.contiguous()
            lib().qrdm_dormqr_dev(0, m, n, min(m, n), ctypes.c_void_p(at.data_ptr()), m,
                                  ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(qr.data_ptr()),
                                  m, None)
            perm = torch.from_numpy(jp.astype(np.int64) - 1).cuda()
            be = (torch.linalg.norm(qr.T - a[:, perm]) / torch.linalg.norm(a)).item()
            d = torch.abs(torch.diagonal(R)).cpu().numpy()
            k = min(m, n)
            ratio = d[:k] / sv[:k]
            res[label] = dict(steps=steps, ms=float(np.median(ts)), backward=be,
                              rr_min=float(ratio.min()), rr_max=float(ratio.max()),
                              mean_block=float(min(m, n) / max(steps, 1)))
        out[fam] = res
        if log:
            log("eval %s: %s" % (fam, res))
    return out


def main(argv=None):
    """python -m qrdm_amd.learned train|eval [options] (on the GPU)."""
    import argparse
    import json
    import time
    import torch
    ap = argparse.ArgumentParser(prog="python -m qrdm_amd.learned")
    ap.add_argument("cmd", choices=["train", "eval", "stats"])
    ap.add_argument("--matrices", type=int, default=6000, help="training matrices in all")
    ap.add_argument("--m", type=int, default=2048)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=THETA_PATH)
    ap.add_argument("--theta", default=THETA_PATH)
    args = ap.parse_args(argv)

    def log(s):
        print(s, flush=True)

    if args.cmd == "train":
        t0 = time.time()
        recs = collect_traces(torch, args.matrices // len(FAMILIES), m=args.m, n=args.n,
                              seed=args.seed, log=log)
        t1 = time.time()
        log(json.dumps(dict(label_stats=label_stats(recs))))
        theta, hist = train(torch, recs[:, :34], recs[:, 34:36], epochs=args.epochs,
                            seed=args.seed, log=log)
        t2 = time.time()
        np.save(args.out, theta.astype(np.float32), allow_pickle=False)
        log(json.dumps(dict(records=int(len(recs)), trace_s=round(t1 - t0, 1),
                            train_s=round(t2 - t1, 1), final_loss=hist[-1][0],
                            heldout_loss=hist[-1][1],
                            default_block=float(recs[:, 36].mean()),
                            label_block=float(recs[:, 37].mean()), out=args.out)))
    elif args.cmd == "stats":
        recs = collect_traces(torch, args.matrices // len(FAMILIES), m=args.m, n=args.n,
                              seed=args.seed, log=log)
        log(json.dumps(label_stats(recs)))
    else:
        net = pack(torch, np.load(args.theta, allow_pickle=False))
        log(json.dumps(evaluate(torch, net, log=log)))


if __name__ == "__main__":
    main()
