"""Writes the golden fixtures of tests/golden/ from the reference compiled
from its own sources (oracle/_ref, built by oracle/build_ref.sh).

Runs in the development container only (the reference does not travel).
Inputs come from tests/golden/inputs.py; the seeds are recorded in each file.

    python tests/golden/make_golden.py [--config2]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

from oracle import pyoracle as po  # noqa: E402
import inputs  # noqa: E402

# (kind, m, n, nb, delta, tau, stop, seed); stop 3 uses eta = 1e-3
SMALL = [
    ("gauss", 40, 30, 8, 0.9, 0.15, 0, 11),
    ("gauss", 25, 60, 16, 0.9, 0.15, 0, 12),
    ("gauss", 64, 64, 64, 0.9, 0.15, 0, 13),
    ("int", 50, 45, 8, 0.9, 0.15, 0, 14),
    ("int", 33, 70, 64, 0.5, 0.3, 0, 15),
    ("zerocols", 48, 40, 16, 0.9, 0.15, 0, 16),
    ("dupcols", 60, 50, 16, 0.9, 0.15, 0, 17),
    ("scaled", 70, 35, 8, 0.9, 0.15, 0, 18),
    ("lowrank", 80, 60, 16, 0.9, 0.15, 3, 19),
    ("lowrank", 70, 90, 8, 0.9, 0.15, 1, 25),
    ("gauss", 90, 70, 3, 0.3, 0.5, 2, 20),
    ("gauss", 1, 7, 64, 0.9, 0.15, 0, 21),
    ("gauss", 9, 1, 64, 0.9, 0.15, 0, 22),
    ("gauss", 130, 100, 64, 1.0, 0.0, 0, 23),
    ("gauss", 100, 130, 32, 0.0, 1.0, 0, 24),
]


def save(name, **kw):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **kw)
    print("wrote", path)


def main():
    if not po.ref_available():
        sys.exit("compiled reference not available: run sh oracle/build_ref.sh")
    # config 1: 1024 x 512 N(0,1), delta 0.9, tau 0.15, nb 64
    a = inputs.gaussian(1024, 512, inputs.CONFIG1_SEED)
    r = po.dgeqrdm(a, impl="ref")
    R = np.triu(r["a"][:512, :])
    save("config1_ref.npz", seed=inputs.CONFIG1_SEED, info=r["info"], jpvt=r["jpvt"],
         ncols=r["ncols"], tau=r["tau"], diag_r=np.diag(R).copy(), r_row0=R[0].copy(),
         r_colnorms=np.linalg.norm(R, axis=0), v_colnorms=np.linalg.norm(np.tril(r["a"], -1), axis=0))
    # small cases with the full factored array
    out = {}
    for i, (kind, m, n, nb, delta, tau, stop, seed) in enumerate(SMALL):
        a = inputs.small_case(kind, m, n, seed)
        r = po.dgeqrdm(a, delta=delta, tau=tau, nb=nb, stop=stop, eta=1e-3, impl="ref")
        out["case%d_params" % i] = np.array([m, n, nb, stop, seed], np.int64)
        out["case%d_kind" % i] = np.array(kind)
        out["case%d_dt" % i] = np.array([delta, tau])
        out["case%d_a" % i] = r["a"]
        out["case%d_jpvt" % i] = r["jpvt"]
        out["case%d_tau" % i] = r["tau"]
        out["case%d_ncols" % i] = r["ncols"]
        out["case%d_info" % i] = np.array(r["info"])
    out["ncases"] = np.array(len(SMALL))
    save("small_ref.npz", **out)
    if "--config2" in sys.argv:
        import time
        t = time.time()
        a = inputs.lowrank_noise(8192, 8192, 2048, inputs.CONFIG2_SEED)
        print("generated config 2 input in %.1fs" % (time.time() - t))
        t = time.time()
        r = po.dgeqrdm(a, impl="ref")
        print("reference dgeqrdm 8192^2 in %.1fs" % (time.time() - t))
        d = np.abs(np.diag(r["a"]))
        # the leading 2048 reflectors (the numerical-rank part): column norms
        # of V, two seeded projections V^T x, and eight whole V columns
        lead = 2048
        V = np.tril(r["a"][:, :lead], -1)
        x = np.random.default_rng(inputs.CONFIG2_SEED + 1).standard_normal((8192, 2))
        vcols = np.arange(0, lead, lead // 8)
        save("config2_ref.npz", seed=inputs.CONFIG2_SEED, info=r["info"], jpvt_lead=r["jpvt"][:lead],
             ncols=r["ncols"], diag_r=d[:2304].copy(), tau_lead=r["tau"][:lead],
             r_row_lead=np.triu(r["a"][:lead, :lead]).sum(axis=0),
             v_colnorms_lead=np.linalg.norm(V, axis=0), v_proj_x=x, v_proj_lead=V.T @ x,
             v_cols_idx=vcols, v_cols=V[:, vcols].copy())


if __name__ == "__main__":
    main()
