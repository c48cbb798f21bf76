"""The batched configuration: 256 independent generalized eigenproblems with n log-uniform in [124, 1204] (seed
1003), H and S per problem as in the n = 1204 configuration, solved by one xtbeig_dev_sygvd_batched call over
concurrent streams, once in plain FP64 and once in FP32 + refinement, with band widths from the autotuner.

Usage: python xtb-diagonalization_amd/batched_bench.py [--count 256] [--slots 8] [--graph] [--check 8]
Prints one JSON line: solves/s and p50 / p99 per-solve latency per precision, the tuned bands, and the FP64 gates of
`--check` problems against scipy.
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--graph", action="store_true")
    ap.add_argument("--reps", type=int, default=2, help="timed passes (the first pass after the warm-up pass)")
    ap.add_argument("--check", type=int, default=8)
    args = ap.parse_args()
    import numpy as np
    import scipy.linalg as sl
    import torch
    import workloads
    import xtbeig

    rng = np.random.default_rng(workloads.SEED_BASE + 3)
    ns = [int(round(np.exp(x))) for x in rng.uniform(np.log(124), np.log(1204), args.count)]
    t0 = time.perf_counter()
    probs = [workloads.generalized_pair(n, workloads.SEED_BASE + 3 + 7 * (q + 1)) for q, n in enumerate(ns)]
    gen_s = time.perf_counter() - t0
    hs = [torch.from_numpy(np.ascontiguousarray(h.T)).cuda() for h, _ in probs]
    ss = [torch.from_numpy(np.ascontiguousarray(s.T)).cuda() for _, s in probs]
    ws = [torch.zeros(n, dtype=torch.float64, device="cuda") for n in ns]
    cs = [torch.zeros((n, n), dtype=torch.float64, device="cuda") for n in ns]
    hnd = xtbeig.Handle(0)
    out = {"config": {"workload": f"{args.count} generalized eigensolves, n log-uniform in [124, 1204], seed 1003",
                      "slots": args.slots, "graph": args.graph, "flops_nominal": float(sum(7.0 * n ** 3 for n in ns))},
           "generate_s": gen_s}
    for name, prec in (("fp64", xtbeig.FP64), ("fp32_refine", xtbeig.FP32_REFINE)):
        tuned = hnd.autotune(prec, 124, 1204)
        hnd.sygvd_batched(hs, ss, ws, cs, prec, slots=args.slots, use_graph=args.graph)  # warm-up pass
        hnd.synchronize()
        best = None
        for _ in range(args.reps):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            lat = hnd.sygvd_batched(hs, ss, ws, cs, prec, slots=args.slots, use_graph=args.graph)
            st = hnd.synchronize()
            wall = time.perf_counter() - t1
            if st != [0, 0, 0, 0]:
                raise RuntimeError(f"status {st}")
            if best is None or wall < best[0]:
                best = (wall, lat)
        wall, lat = best
        gates = []
        for q in range(min(args.check, args.count)):
            h, s = probs[q]
            g = workloads.gates(h, s, ws[q].cpu().numpy(), cs[q].cpu().numpy().T, sl.eigh(h, s, eigvals_only=True))
            gates.append(workloads.passes(g))
        out[name] = {"solves_per_s": args.count / wall, "wall_s": wall,
                     "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                     "nominal_TFLOPs": out["config"]["flops_nominal"] / wall / 1e12,
                     "bands": {str(b0): b for b0, b, _ in tuned}, "checked_pass": all(gates), "checked": len(gates)}
        print(json.dumps({name: out[name]}), file=sys.stderr, flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
