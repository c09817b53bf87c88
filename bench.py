"""bench.py -- DM-QR time and TF/s on one MI355X.

A "step" is one full DM-QR factorisation (dgeqrdm semantics, FP64) of the
8192 x 8192 low-rank-plus-noise matrix A = G1 G2 + 1e-10 N (G1 8192 x 2048,
G2 2048 x 8192, N(0,1) entries), delta = 0.9, tau = 0.15, nb = 64, with the
matrix resident in HBM.  Each timed step starts from the same pristine copy
(restored on the device before the step's clock starts).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--m M] [--n N]
                    [--rank R] [--no-cpu-baseline] [--no-configs]

Prints ONE JSON line (rank 0).  value = whole-factorisation TF/s with
F = 2 m n^2 - 2 n^3 / 3.  roofline = the trailing update (k_fused, the
merged pass over C, and k_wsum_norm), its kernels timed with HIP events on
the library's launch stream during separate event-timed steps.
cpu_baseline = the reference compiled from its own sources (oracle/_ref,
OpenBLAS) on the same workload, on this host's cores.  configs = the other
workloads of the headline metric (1024 x 512 beside the CPU reference and
dgeqp3, the 16384^2 graded matrix) and the drop-in host-pointer time.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("DM-QR time & TF/s, 8192² FP64; mixed/learned vs plain at "
          "||AP-QR||/||A||≤1e-12")
FP64_MFMA_PEAK_TFS = 78.6
HBM_PEAK_GBS = 8000.0
# HBM bytes per k_fused launch (FETCH_SIZE + WRITE_SIZE, each scaled by the
# counter's reading for a transfer of known size in the same run: the initial
# norm sweep reads the matrix once, the restore copy writes it once), averaged
# over the launches of the workload (m, n, rank, nb); measured with
# rocprofv3 --pmc, summary in profiles/k_fused_8192_pmc_traffic.md.
FUSED_TRAFFIC = {(8192, 8192, 2048, 64): 366.5e6}


def fused_algorithmic_bytes(m, n, ncols):
    """Mean algorithmic HBM bytes of one k_fused launch over a factorisation
    whose DM steps reduced ncols[it] columns: C (m_sub x (n - c0)) read and
    written once, the two V blocks (m_sub x 64 each) and the previous W
    (64 x (n - c0)) read once; c0 = j + k."""
    j, tot, cnt = 0, 0.0, 0
    for k in ncols:
        k = int(k)
        if k <= 0:
            break
        msub, nct = m - j, n - j - k
        if nct > 0 and msub > 0:
            tot += 16.0 * msub * nct + 16.0 * msub * 64 + 8.0 * 64 * nct
            cnt += 1
        j += k
    return tot / max(cnt, 1)


def flops(m, n):
    if m >= n:
        return 2.0 * m * n * n - 2.0 * n ** 3 / 3.0
    return 2.0 * n * m * m - 2.0 * m ** 3 / 3.0


def make_input(torch, m, n, r, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    # (n, m) row-major = the column-major m x n matrix
    g2t = torch.randn(n, r, dtype=torch.float64, device="cuda", generator=g)
    g1t = torch.randn(r, m, dtype=torch.float64, device="cuda", generator=g)
    a = g2t @ g1t
    a += 1e-10 * torch.randn(n, m, dtype=torch.float64, device="cuda", generator=g)
    return a


def host_cpus():
    """The CPUs this process may run on and what lscpu says about them."""
    info = {"cpus_available": len(os.sched_getaffinity(0))}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name")
        info["sockets"] = int(kv.get("Socket(s)", "0") or 0)
        info["cores_per_socket"] = int(kv.get("Core(s) per socket", "0") or 0)
        info["threads_per_core"] = int(kv.get("Thread(s) per core", "0") or 0)
    except Exception:
        pass
    return info


def median_time(fn, runs, warmup=1):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2], ts


def cpu_baseline(n_sample=8192, rank_sample=2048, runs=5):
    """The compiled reference (oracle/_ref) on the benched workload: one
    warm-up, then the median of `runs` factorisations on all the CPUs this
    process may use; dgeqp3 of the same library beside it (median of 3); a
    one-thread row on 4096^2 of the same family, extrapolated to the benched
    size by the n^3 flop count."""
    try:
        from oracle import pyoracle as po
        sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
        import inputs
    except Exception as e:  # pragma: no cover
        return {"error": "oracle unavailable: %s" % e}
    if not po.ref_available():
        return {"error": "compiled reference (oracle/_ref) not present"}
    cpus = host_cpus()
    cores = max(1, cpus["cpus_available"])
    po.lib().oracle_ref_set_threads(cores)
    a = inputs.lowrank_noise(n_sample, n_sample, rank_sample, 77)
    info = []
    t_dm, ts_dm = median_time(lambda: info.append(po.dgeqrdm(a, impl="ref")["info"]), runs)
    f = flops(n_sample, n_sample)
    # dgeqp3 (half its flops in matrix-vector products) and the one-thread
    # reference: one run each on 4096^2 of the same family, scaled to the
    # benched size by the flop count, to keep the host's share of the run short
    n1 = 4096
    a1 = inputs.lowrank_noise(n1, n1, n1 // 4, 78)
    t_qp3_1, _ = median_time(lambda: info.append(po.dgeqp3_ref(a1)["info"]), 1, warmup=0)
    t_qp3 = t_qp3_1 * f / flops(n1, n1)
    po.lib().oracle_ref_set_threads(1)
    t1, _ = median_time(lambda: po.dgeqrdm(a1, impl="ref"), 1, warmup=0)
    po.lib().oracle_ref_set_threads(cores)
    one = {"value": round(flops(n1, n1) / t1 / 1e12, 5), "unit": "TF/s", "cores": 1,
           "seconds": round(t1, 3),
           "extrapolated_seconds_%dx%d" % (n_sample, n_sample): round(t1 * f / flops(n1, n1), 1),
           "sample": "reference dgeqrdm, 1 thread, %dx%d low-rank(%d)+1e-10 noise, one run; "
                     "extrapolated to %dx%d by the flop count" % (n1, n1, n1 // 4, n_sample,
                                                                  n_sample)}
    return {"value": round(f / t_dm / 1e12, 5), "unit": "TF/s", "cores": cores,
            "host": cpus, "one_thread": one, "kind": "reference",
            "sample": "reference dgeqrdm (compiled from its sources, OpenBLAS 0.3.28, "
                      "%d threads) on the benched workload (%dx%d low-rank(%d)+1e-10 noise), "
                      "median of %d after one warm-up: %.2f s; LAPACKE_dgeqp3 (same library "
                      "and threads) on %dx%d of the same family, one run, %.2f s, scaled to "
                      "%.1f s (%.5f TF/s)"
                      % (cores, n_sample, n_sample, rank_sample, runs, t_dm, n1, n1, t_qp3_1,
                         t_qp3, f / t_qp3 / 1e12),
            "info": int(sum(abs(i) for i in info)),
            "seconds": round(t_dm, 3), "seconds_all": [round(t, 3) for t in ts_dm],
            "dgeqp3_seconds_scaled": round(t_qp3, 3)}


def config1(torch, qrdm_amd, runs=5):
    """1024 x 512 N(0,1) (the seeded fixture input): the GPU through the
    drop-in host-pointer dgeqrdm and through the device-resident form, beside
    the compiled reference's dgeqrdm and LAPACKE_dgeqp3 on the host (each a
    median of `runs` after one warm-up)."""
    import ctypes
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import inputs
    m, n = 1024, 512
    a = inputs.gaussian(m, n, inputs.CONFIG1_SEED)
    L = qrdm_amd.lib()
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    th = np.array([0.9, 0.15, 0.0])

    def host_call():
        af = a.copy(order="F")
        jp = np.zeros(n, np.int32)
        nc = np.zeros(n, np.int32)
        t = np.zeros(n)
        info = L.dgeqrdm(102, m, n, af.ctypes.data_as(dp), m, jp.ctypes.data_as(ip),
                         t.ctypes.data_as(dp), nc.ctypes.data_as(ip), th.ctypes.data_as(dp), 64)
        if info != 0:
            raise RuntimeError("dgeqrdm info %d" % info)
        return jp
    t_host, _ = median_time(host_call, runs)
    jp_gpu = host_call()
    pristine = torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
    work = torch.empty_like(pristine)
    tau = torch.zeros(n, dtype=torch.float64, device="cuda")

    def dev_call():
        work.copy_(pristine)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        jp = np.zeros(n, np.int32)
        nc = np.zeros(n, np.int32)
        info, _ = qrdm_amd.dgeqrdm_dev(work, m, tau, jp, nc, th)
        torch.cuda.synchronize()
        if info != 0:
            raise RuntimeError("dgeqrdm_dev info %d" % info)
        return time.perf_counter() - t0
    dev_call()
    t_dev = sorted(dev_call() for _ in range(runs))[runs // 2]
    out = {"workload": "dgeqrdm 1024x512 N(0,1), delta=0.9 tau=0.15 nb=64 (seeded fixture input)",
           "flops": flops(m, n),
           "gpu_host_pointer_ms": round(1000 * t_host, 3),
           "gpu_device_resident_ms": round(1000 * t_dev, 3),
           "gpu_device_resident_tfs": round(flops(m, n) / t_dev / 1e12, 4)}
    try:
        from oracle import pyoracle as po
        if po.ref_available():
            cores = max(1, len(os.sched_getaffinity(0)))
            po.lib().oracle_ref_set_threads(cores)
            ref = {}
            t_ref, _ = median_time(lambda: ref.update(po.dgeqrdm(a, impl="ref")), runs)
            t_qp3, _ = median_time(lambda: po.dgeqp3_ref(a), runs)
            out.update({"cpu_reference_dgeqrdm_ms": round(1000 * t_ref, 3),
                        "cpu_dgeqp3_ms": round(1000 * t_qp3, 3), "cpu_threads": cores,
                        "jpvt_identical": bool(np.array_equal(ref["jpvt"], jp_gpu))})
    except Exception as e:  # pragma: no cover
        out["cpu_error"] = str(e)
    return out


def config3(torch, qrdm_amd, k=16384, runs=2):
    """16384^2 graded spectrum A = U diag(s) V^T, s_i = 10^(-8 i / n), U and V
    from the QR of seeded Gaussians (built on the device): the FP64 path's time
    (median of `runs` after one warm-up) and its backward error."""
    import ctypes
    import numpy as np
    g = torch.Generator(device="cuda")
    g.manual_seed(2026)
    L = qrdm_amd.lib()

    def orthogonal():
        # Q of the unpivoted QR (the library's dgeqrf and dormqr) of a Gaussian
        x = torch.randn(k, k, dtype=torch.float64, device="cuda", generator=g)
        t = torch.zeros(k, dtype=torch.float64, device="cuda")
        if L.qrdm_dgeqrf_dev(k, k, ctypes.c_void_p(x.data_ptr()), k,
                             ctypes.c_void_p(t.data_ptr()), None) != 0:
            raise RuntimeError("dgeqrf_dev")
        q = torch.eye(k, dtype=torch.float64, device="cuda")
        if L.qrdm_dormqr_dev(0, k, k, k, ctypes.c_void_p(x.data_ptr()), k,
                             ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(q.data_ptr()), k,
                             None) != 0:
            raise RuntimeError("dormqr_dev")
        torch.cuda.synchronize()
        return q.T  # (rows, cols) view of the column-major Q
    u = orthogonal()
    v = orthogonal()
    sv = 10.0 ** (-8.0 * torch.arange(k, dtype=torch.float64, device="cuda") / k)
    pristine = ((v * sv) @ u.T).contiguous()  # (n, m) row-major = U diag(s) V^T column-major
    del u, v
    work = torch.empty_like(pristine)
    tau = torch.zeros(k, dtype=torch.float64, device="cuda")
    th = np.array([0.9, 0.15, 0.0])
    jp = np.zeros(k, np.int32)

    def run():
        work.copy_(pristine)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        jp[:] = 0
        nc = np.zeros(k, np.int32)
        info, _ = qrdm_amd.dgeqrdm_dev(work, k, tau, jp, nc, th)
        torch.cuda.synchronize()
        if info != 0:
            raise RuntimeError("dgeqrdm_dev info %d" % info)
        return time.perf_counter() - t0, int(np.count_nonzero(nc))
    run()
    ts = sorted(run() for _ in range(runs))
    t, steps = ts[len(ts) // 2]
    stream = torch.cuda.current_stream().cuda_stream
    be = backward_error(torch, qrdm_amd, work, pristine, tau, jp, stream)
    return {"workload": "dgeqrdm %dx%d graded s_i = 10^(-8 i/n), FP64" % (k, k),
            "ms": round(1000 * t, 2), "tfs": round(flops(k, k) / t / 1e12, 3), "dm_steps": steps,
            "backward_error": float("%.3g" % be)}


def host_pointer_8192(torch, qrdm_amd, pristine, m, n):
    """The drop-in dgeqrdm on host memory at the benched size: the matrix
    copied to the device and back inside the call (PCIe-inclusive)."""
    import ctypes
    import numpy as np
    a = np.asfortranarray(pristine.cpu().numpy().T)
    L = qrdm_amd.lib()
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    th = np.array([0.9, 0.15, 0.0])

    def call():
        af = a.copy(order="F")
        jp = np.zeros(n, np.int32)
        nc = np.zeros(n, np.int32)
        t = np.zeros(min(m, n))
        t0 = time.perf_counter()
        info = L.dgeqrdm(102, m, n, af.ctypes.data_as(dp), m, jp.ctypes.data_as(ip),
                         t.ctypes.data_as(dp), nc.ctypes.data_as(ip), th.ctypes.data_as(dp), 64)
        dt = time.perf_counter() - t0
        if info != 0:
            raise RuntimeError("dgeqrdm info %d" % info)
        return dt
    call()
    ts = sorted(call() for _ in range(3))
    return {"workload": "dgeqrdm %dx%d from host memory (pageable numpy array, copies in and "
                        "out inside the call)" % (m, n),
            "ms": round(1000 * ts[1], 2), "tfs": round(flops(m, n) / ts[1] / 1e12, 3)}


def progress(msg):
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def batch_shapes(count, seed, mmin=64, mmax=1024, nmin=32, nmax=512):
    """Uneven batch shapes: m uniform in [mmin, mmax], n uniform in
    [nmin, min(m, nmax)]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ms = rng.integers(mmin, mmax + 1, count)
    return [(int(m), int(rng.integers(nmin, min(int(m), nmax) + 1))) for m in ms]


def batch_fill(torch, plan, seed, lo=-10.0):
    """The packed batch tensor of plan and the spectra: matrix i = U diag(s) V^T, s sorted
    log-uniform over [10^lo, 1], U (m x k) and V (n x k) with orthonormal
    columns from the QR of seeded Gaussians (grouped by row count so that
    each group is one batched QR on the device)."""
    import numpy as np
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    rng = np.random.default_rng(seed + 1)
    shapes = plan.shapes
    ks = [min(m, n) for m, n in shapes]
    facs = [[None, None] for _ in shapes]
    for side in (0, 1):
        groups = {}
        for i, (m, n) in enumerate(shapes):
            groups.setdefault((m, n)[side], []).append(i)
        for rows in sorted(groups):
            idx = groups[rows]
            kmax = max(ks[i] for i in idx)
            g = torch.randn(len(idx), rows, kmax, generator=gen, device="cuda",
                            dtype=torch.float64)
            q = torch.linalg.qr(g).Q
            for b, i in enumerate(idx):
                facs[i][side] = q[b, :, :ks[i]]
    buf = torch.empty(plan.a_size, dtype=torch.float64, device="cuda")
    spectra = []
    for i, (m, n) in enumerate(shapes):
        k = ks[i]
        s = np.sort(10.0 ** rng.uniform(lo, 0.0, k))[::-1].copy()
        spectra.append(s)
        st = torch.from_numpy(s).cuda()
        a = (facs[i][0] * st) @ facs[i][1].T  # (m, n)
        o = int(plan.a_off[i])
        buf[o:o + m * n] = a.T.contiguous().view(-1)
    return buf, spectra


def backward_error(torch, qrdm_amd, work, pristine, tau, jp, stream):
    """||A P - Q R||_F / ||A||_F of the factorisation held in work (R and the
    reflectors, column-major m x n in an (n, m) tensor), Q R formed by the
    library's dormqr on the device."""
    import ctypes
    n, m = work.shape
    R = torch.triu(work.T)
    qr = R.T.contiguous()
    info = qrdm_amd.lib().qrdm_dormqr_dev(0, m, n, min(m, n), ctypes.c_void_p(work.data_ptr()),
                                          m, ctypes.c_void_p(tau.data_ptr()),
                                          ctypes.c_void_p(qr.data_ptr()), m,
                                          ctypes.c_void_p(stream))
    if info != 0:
        raise RuntimeError("dormqr_dev info %d" % info)
    A = pristine.T
    perm = torch.from_numpy(jp.astype("int64") - 1).to(work.device)
    return (torch.linalg.norm(qr.T - A[:, perm]) / torch.linalg.norm(A)).item()


def trailing_forms(torch, qrdm_amd, args, timed, work, pristine, tau, last_jp, stream, ms_plain,
                   which):
    """The plain FP64 path beside the FP32 and split-BF16x3 trailing updates
    (scheduled: FP64 until the remaining block's norm bound allows the lower
    precision within 1e-12; _all: the lower precision in every step) on the
    same matrix: time per factorisation (same timing as the headline, no
    events) and the backward error of one factorisation of each."""
    m, n = args.m, args.n
    out = {}
    for form in ["fp64"] + list(which):
        ms_f = ms_plain
        if form != "fp64":
            timed(1, form)  # warm-up of the form's kernels
            ms_f, _ = timed(args.steps, form)
        # one more factorisation of this form, left in work, for its backward error
        timed(1, form)
        be = backward_error(torch, qrdm_amd, work, pristine, tau, last_jp, stream)
        out[form] = {"ms_per_step": round(ms_f, 3),
                     "tfs": round(flops(m, n) / (ms_f / 1000.0) / 1e12, 4),
                     "speedup_vs_fp64": round(ms_plain / ms_f, 4),
                     "backward_error": float("%.4g" % be),
                     "meets_1e-12": bool(be <= 1e-12)}
    return out


def batched_form(torch, qrdm_amd, steps, loop_count, autotune=True):
    """4096 uneven matrices (m in [64, 1024], n in [32, min(m, 512)], spectra
    log-uniform over [1e-10, 1]) in one persistent launch (4 waves per
    workgroup, nb 64, the panel form autotuned, replayed from a hipGraph),
    beside the column-by-column panel form and a loop of the plain path."""
    shapes = batch_shapes(4096, 4)
    plan = qrdm_amd.BatchPlan(shapes, nb=64)
    src, _ = batch_fill(torch, plan, 4)
    q = plan.autotune(src) if autotune else plan.query()
    rows = q["panel_rows"]
    work = torch.empty_like(src)
    tau = torch.zeros(plan.tau_size, dtype=torch.float64, device="cuda")
    jp = torch.zeros(plan.col_size, dtype=torch.int32, device="cuda")
    nc = torch.zeros(plan.col_size, dtype=torch.int32, device="cuda")
    info = torch.zeros(len(shapes), dtype=torch.int32, device="cuda")
    total = 0.0
    for rep in range(steps + 1):
        work.copy_(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.run(work, tau, jp, nc, info, graph=True)
        torch.cuda.synchronize()
        if rep > 0:
            total += time.perf_counter() - t0
    ms = 1000.0 * total / steps
    if int((info != 0).sum().item()) != 0:
        raise RuntimeError("batched: non-zero info")
    # the column-by-column panel form, for reference
    plan.set_panel_rows(0)
    ms_cbc = []
    for rep in range(2):
        work.copy_(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.run(work, tau, jp, nc, info, graph=True)
        torch.cuda.synchronize()
        ms_cbc.append(1000.0 * (time.perf_counter() - t0))
    plan.set_panel_rows(rows)
    import numpy as np
    work.copy_(src)
    tau2 = torch.zeros_like(tau)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, (m, n) in enumerate(shapes[:loop_count]):
        o, t = int(plan.a_off[i]), int(plan.tau_off[i])
        jv = np.zeros(n, np.int32)
        nv = np.zeros(n, np.int32)
        r, _ = qrdm_amd.dgeqrdm_dev(work[o:o + m * n].view(n, m), m, tau2[t:t + min(m, n)], jv,
                                    nv, np.array([0.9, 0.15, 0.0]))
        if r != 0:
            raise RuntimeError("loop: info %d" % r)
    torch.cuda.synchronize()
    loop_ms = 1000.0 * (time.perf_counter() - t0)
    fl = sum(flops(max(m, n), min(m, n)) for m, n in shapes)
    plan.close()
    return {"workload": "4096 uneven FP64 matrices, m in [64,1024], n in [32,min(m,512)], "
                        "U diag(s) V^T with s log-uniform over [1e-10,1], nb=64",
            "ms_per_batch": round(ms, 2), "tfs": round(fl / (ms / 1000.0) / 1e12, 4),
            "ms_column_by_column_panel": round(ms_cbc[-1], 2),
            "flops": fl, "per_matrix_loop_matrices": min(loop_count, len(shapes)),
            "per_matrix_loop_ms": round(loop_ms, 1),
            "speedup_vs_loop": (round(loop_ms / ms, 2) if loop_count >= len(shapes) else None),
            "launch": q}


def learned_form(torch):
    """The committed (tau, delta) network against the defaults on one
    8192 x 4096 matrix of each training family."""
    from qrdm_amd import learned
    net = learned.load_policy(torch)
    res = learned.evaluate(torch, net, reps=2)
    out = {}
    for fam, r in res.items():
        d, l = r["default"], r["learned"]
        out[fam] = {"steps": [d["steps"], l["steps"]], "ms": [round(d["ms"], 3), round(l["ms"], 3)],
                    "backward": [float("%.3g" % d["backward"]), float("%.3g" % l["backward"])],
                    "min_Rkk_over_sigma": [float("%.3g" % d["rr_min"]),
                                           float("%.3g" % l["rr_min"])]}
    return {"workload": "8192x4096 FP64, one matrix per family (geometric, step, Kahan-type); "
                        "pairs are [defaults, learned]", "families": out}


def form_in_child(key, batch_loop, limit=400):
    """One added form measured by a child bench.py process (its JSON on
    stdout); a failure there is reported in the line and leaves the headline
    measurement alone."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--form-child", key,
           "--batch-loop", str(batch_loop)]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=limit)
    except subprocess.TimeoutExpired:
        return {"error": "child process over %d s" % limit}
    if r.returncode != 0:
        return {"error": "child process exit %d: %s" % (r.returncode, r.stderr[-300:])}
    try:
        return json.loads(r.stdout.strip().splitlines()[-1])
    except Exception as e:
        return {"error": "child output: %s" % e}


def run_form_child(key, batch_loop):
    import torch
    import qrdm_amd
    torch.cuda.set_device(0)
    if key == "batched":
        out = batched_form(torch, qrdm_amd, 3, batch_loop, autotune=True)
    elif key == "learned":
        out = learned_form(torch)
    else:
        raise SystemExit("unknown form %s" % key)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--m", type=int, default=8192)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--rank", type=int, default=2048)
    ap.add_argument("--nb", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 1024x512, 16384^2 and host-pointer runs")
    ap.add_argument("--forms", default="fp32,bf16x3,fp32_all,batched,learned",
                    help="comma list of the added forms to measure beside the plain path "
                         "(fp32, bf16x3: the scheduled reduced-precision trailing updates; "
                         "fp32_all, bf16x3_all: the reduced precision in every step; "
                         "batched, learned; empty for none).  batched and "
                         "learned run in child processes (a failure there is reported in the "
                         "line)")
    ap.add_argument("--form-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--batch-loop", type=int, default=4096,
                    help="matrices in the per-matrix loop the batched form is compared with")
    args = ap.parse_args()
    if args.form_child:
        run_form_child(args.form_child, args.batch_loop)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(the multi-GPU design is described, not built)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import qrdm_amd

    m, n = args.m, args.n
    torch.cuda.set_device(0)
    pristine = make_input(torch, m, n, args.rank, 1234)
    work = torch.empty_like(pristine)
    tau = torch.zeros(min(m, n), dtype=torch.float64, device="cuda")
    thres = np.array([0.9, 0.15, 0.0])
    stream = torch.cuda.current_stream().cuda_stream

    last_jp = np.zeros(n, np.int32)

    def one_step(trailing="fp64"):
        jp = np.zeros(n, np.int32)
        nc = np.zeros(n, np.int32)
        info, _ = qrdm_amd.dgeqrdm_dev(work, m, tau, jp, nc, thres, nb=args.nb, stream=stream,
                                       trailing=trailing)
        if info != 0:
            raise RuntimeError("dgeqrdm_dev info %d: %s" % (info, qrdm_amd.last_error()))
        last_jp[:] = jp
        return nc

    for _ in range(args.warmup):
        work.copy_(pristine)
        torch.cuda.synchronize()
        one_step()
    torch.cuda.synchronize()

    def timed(steps, trailing="fp64"):
        total, nc = 0.0, None
        for _ in range(steps):
            work.copy_(pristine)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            nc = one_step(trailing)
            torch.cuda.synchronize()
            total += time.perf_counter() - t0
        return 1000.0 * total / steps, nc

    # timed steps, no events on the stream
    qrdm_amd.profile(enable=False, reset=True)
    ms, last_nc = timed(args.steps)
    progress("timed steps: %.3f ms per factorisation" % ms)
    nsteps_dm = int(np.count_nonzero(last_nc))
    tfs = flops(m, n) / (ms / 1000.0) / 1e12
    # the same steps again with HIP events around the trailing update's two
    # kernels (each event record idles the stream for a few microseconds, so
    # these steps give the kernels' average durations, not the headline time)
    qrdm_amd.profile(enable=True, reset=True, classes=["larfb_w", "larfb_sum"])
    ms_ev, _ = timed(args.steps)
    prof = qrdm_amd.profile(enable=False, reset=True)

    # one more (untimed) factorisation with every kernel class timed, for the
    # per-class breakdown
    work.copy_(pristine)
    torch.cuda.synchronize()
    qrdm_amd.profile(enable=True, reset=True)
    one_step()
    torch.cuda.synchronize()
    prof_all = qrdm_amd.profile(enable=False, reset=True)

    # the trailing update: k_fused (C -= V_prev W_prev with the next step's
    # V^T C, one read and one write of C) and k_wsum_norm (W = T^T sum, the
    # R12 rows and the norm downdate); FP64 MFMA bound
    ms_f, launches_f, flops_f = prof["larfb_w"]
    ms_s, launches_s, flops_s = prof["larfb_sum"]
    tot_ms = ms_f + ms_s
    achieved = (flops_f + flops_s) / (tot_ms / 1000.0) / 1e12 if tot_ms > 0 else 0.0
    alg_bytes = fused_algorithmic_bytes(m, n, last_nc)
    bytes_f = alg_bytes * launches_f
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_MFMA_PEAK_TFS,
                "unit": "TF/s", "frac": round(achieved / FP64_MFMA_PEAK_TFS, 4),
                "traffic": FUSED_TRAFFIC.get((m, n, args.rank, args.nb)),
                "algorithmic_bytes": round(alg_bytes),
                "kernel": "trailing update: k_fused (%d launches, avg %.4f ms) + k_wsum_norm "
                          "(%d launches, avg %.4f ms); event-timed steps %.3f ms per "
                          "factorisation" % (launches_f, ms_f / max(launches_f, 1), launches_s,
                                             ms_s / max(launches_s, 1), ms_ev),
                "kernels": {
                    "k_fused": {"ms": round(ms_f, 3), "launches": int(launches_f),
                                "tflops": round(flops_f / 1e12, 4),
                                "achieved_tfs": round(flops_f / (ms_f / 1000.0) / 1e12, 3)
                                if ms_f > 0 else None,
                                "achieved_GBs": round(bytes_f / (ms_f / 1000.0) / 1e9, 1)
                                if ms_f > 0 else None},
                    "k_wsum_norm": {"ms": round(ms_s, 3), "launches": int(launches_s),
                                    "tflops": round(flops_s / 1e12, 4)}}}
    others = {}
    for c in ("larfb_w", "larfb_sum", "select", "panel", "permute", "init_norms"):
        t, k, w = prof_all[c]
        entry = {"ms": round(t, 3), "launches": int(k)}
        if c in ("select", "permute", "init_norms") and t > 0:
            entry["achieved_GBs"] = round(w / (t / 1000.0) / 1e9, 1)
            entry["frac_hbm"] = round(w / (t / 1000.0) / 1e9 / HBM_PEAK_GBS, 4)
        others[{"larfb_w": "k_fused", "larfb_sum": "k_wsum_norm"}.get(c, c)] = entry
    line = {"metric": METRIC, "value": round(tfs, 4), "unit": "TF/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "dgeqrdm %dx%d low-rank(%d)+1e-10 noise, delta=0.9 "
                                   "tau=0.15 nb=%d, matrix resident in HBM"
                                   % (m, n, args.rank, args.nb),
                       "m": m, "n": n, "rank": args.rank, "nb": args.nb,
                       "dm_steps": nsteps_dm,
                       "kernel_classes_one_factorisation": others},
            "roofline": roofline}
    progress("profiled steps done")
    forms = [f for f in args.forms.split(",") if f]
    line["forms"] = {}
    if forms:  # the plain path's own backward error beside the forms
        line["forms"] = trailing_forms(torch, qrdm_amd, args, timed, work, pristine, tau,
                                       last_jp, stream, ms,
                                       [f for f in forms if f in ("fp32", "bf16x3", "fp32_all",
                                                                   "bf16x3_all")])
    for key in ("batched", "learned"):
        if key in forms:
            line["forms"][key] = form_in_child(key, args.batch_loop)
            progress("form %s done" % key)
    progress("forms done")
    if not args.no_configs:
        line["configs"] = {"host_pointer_%dx%d" % (m, n): host_pointer_8192(torch, qrdm_amd,
                                                                            pristine, m, n)}
        progress("host-pointer run done")
        line["configs"]["config1_1024x512"] = config1(torch, qrdm_amd)
        progress("config 1 done")
        torch.cuda.empty_cache()
        line["configs"]["config3_16384_fp64"] = config3(torch, qrdm_amd)
        progress("config 3 done")
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
        progress("cpu baseline done")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
