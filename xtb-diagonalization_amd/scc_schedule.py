"""The SCC-like precision sequence of the mixed-precision configuration, run through the library's schedule API.

H_k = H_0 + D_k, k = 0 .. 19, D_k symmetric N(0, 1) scaled to ||D_k||_F = 1e-1 (1e-6)^(k / 19); the schedule runs a
cycle in FP32 while ||D_k||_F >= fp32_rmsq (the proxy for the rms charge change the reference switches on,
scripts/fp32seed.py:15-17) and in FP64 from then on. Every cycle's eigenvalues and eigenvectors are checked against
the FP64 host-LAPACK oracle (scipy.linalg.eigh) with the contract's three gates.

Usage: python xtb-diagonalization_amd/scc_schedule.py [--n 1788] [--band 32] [--rmsq 1e-3] [--refine 1]
Prints one JSON line per cycle and a summary line.
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def sequence(h0, steps, seed):
    import numpy as np
    rng = np.random.default_rng(seed)
    n = h0.shape[0]
    for k in range(steps):
        d = rng.standard_normal((n, n))
        d = 0.5 * (d + d.T)
        norm = 1e-1 * (1e-6) ** (k / (steps - 1))
        d *= norm / np.linalg.norm(d)
        yield k, norm, np.asfortranarray(h0 + d)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1788)
    ap.add_argument("--band", type=int, default=32)
    ap.add_argument("--rmsq", type=float, default=1e-3)
    ap.add_argument("--refine", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    args = ap.parse_args()
    import numpy as np
    import scipy.linalg as sl
    import torch
    import workloads
    import xtbeig

    seed = workloads.SEED_BASE + 2
    h0, s = workloads.generalized_pair(args.n, seed)
    hnd = xtbeig.Handle(0)
    hnd.set_tuning(args.band, 2)
    hnd.set_scc_precision(args.rmsq, 160, args.refine)
    hnd.scc_begin()
    sd = torch.from_numpy(np.ascontiguousarray(s.T)).cuda()
    w = torch.zeros(args.n, dtype=torch.float64, device="cuda")
    c = torch.zeros((args.n, args.n), dtype=torch.float64, device="cuda")
    rows = []
    for k, norm, hk in sequence(h0, args.steps, seed + 100):
        hd = torch.from_numpy(np.ascontiguousarray(hk.T)).cuda()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        used = hnd.scc_sygvd(hd, sd, w, c, norm)
        ms = (time.perf_counter() - t0) * 1e3
        wref = sl.eigh(hk, s, eigvals_only=True)
        g = workloads.gates(hk, s, w.cpu().numpy(), c.cpu().numpy().T, wref)
        row = {"step": k, "norm_D": norm, "precision": "fp32_refine" if used == xtbeig.FP32_REFINE else "fp64",
               "ms": ms, "eig_err": g["eig_err"], "residual": g["residual"], "orth": g["orth"],
               "pass": workloads.passes(g)}
        rows.append(row)
        print(json.dumps(row), flush=True)
    settings, counters, rc = hnd.get_scc_precision()
    print(json.dumps({"n": args.n, "band": args.band, "fp32_rmsq": settings[0], "fp32_maxiter": settings[1],
                      "fp32_refine": settings[2], **counters, "all_pass": all(r["pass"] for r in rows)}))


if __name__ == "__main__":
    main()
