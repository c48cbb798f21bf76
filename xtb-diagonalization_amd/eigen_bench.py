"""Dump-driven eigensolver bench, after the reference's eigen-bench (Sources/EigenBench/main.swift:148-420): every
variant on each input, stage times, accuracy, and the FNV-1a hash of each two-stage variant's tridiagonal.

Inputs: HS.bin dumps (read by the library's C reader; H' = X^T H X formed in FP64 on the GPU, as the reference forms
it with dpotrf, dtrtri and two dtrmm, main.swift:164-172) or "random:N", the reference's random symmetric N x N
matrix (entries 2 drand48() - 1 after srand48(N), filled column by column over the upper triangle, main.swift:153-159;
drand48 restated below from its POSIX definition).

Variants, each in FP64 (f64) and FP32 (f32):
  1stage-<p>         one-stage reduction (xtb_[ds]sytrd_upper's device twin) + divide and conquer + dormtr on m columns
  2stage-<p>-b<b>    dense -> band b + bulge chase + divide and conquer + bulge and band back-transformations on m
                     columns, for every band b of --bands
  syevd-<p>          the library's standard solve (the fused two-stage pipeline)
Per variant (median of --reps runs after one warm-up, HIP events per stage): stage times in ms, then the accuracy of
the first m = round(fraction n) columns against the host LAPACK's FP64 eigenvalues, as main.swift:233-251 reports
it: max |w - w_ref|, residual max |H' Z - Z diag(w)| / max |w_ref|, orthogonality max |Z^T Z - I|; two-stage
variants also print the FNV-1a hash of their (d, e) bits in their own precision (main.swift:253-266) and whether
every run gave the same hash.

Usage: python xtb-diagonalization_amd/eigen_bench.py [inputs ...] [--fraction 0.5] [--bands 8,16,32]
       [--precisions f64,f32] [--reps 3]
Prints one JSON line per (input, variant).
"""
import argparse
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

DRAND48_A = 0x5DEECE66D
DRAND48_C = 0xB
DRAND48_M = (1 << 48) - 1


def drand48_stream(seed, count):
    """count values of POSIX drand48 after srand48(seed): X_0 = (seed << 16) | 0x330E (seed's low 32 bits),
    X_{k+1} = (0x5DEECE66D X_k + 11) mod 2^48, value X_{k+1} / 2^48."""
    x = ((seed & 0xFFFFFFFF) << 16) | 0x330E
    out = np.empty(count)
    for k in range(count):
        x = (DRAND48_A * x + DRAND48_C) & DRAND48_M
        out[k] = x / float(1 << 48)
    return out


def random_matrix(n):
    """The reference's "random:N" matrix (main.swift:153-159)."""
    vals = drand48_stream(n, n * (n + 1) // 2)
    h = np.zeros((n, n))
    k = 0
    for j in range(n):
        for i in range(j + 1):
            x = 2.0 * vals[k] - 1.0
            k += 1
            h[i, j] = x
            h[j, i] = x
    return h


def load(torch, xtbeig, hnd, spec):
    """(name, H') for one input, H' symmetric FP64 (numpy, column-major)."""
    if spec.startswith("random:"):
        return spec, random_matrix(int(spec[7:]))
    h, s, _, _ = xtbeig.hsbin_read(spec)
    with torch.cuda.stream(torch.cuda.ExternalStream(hnd.stream_ptr())):
        return os.path.basename(spec), _congruence(torch, hnd, h, s, spec)


def _congruence(torch, hnd, h, s, spec):
    n = h.shape[0]
    hd = torch.from_numpy(np.ascontiguousarray(h.T)).cuda()
    sd = torch.from_numpy(np.ascontiguousarray(s.T)).cuda()
    u = torch.zeros_like(sd)
    x = torch.zeros_like(sd)
    scopy = sd.clone()
    hnd.potrf_inverse(scopy, u, x)
    hx = torch.zeros_like(hd)
    hp = torch.zeros_like(hd)
    hnd.gemm("N", "N", n, n, n, 1.0, hd, n, x, n, 0.0, hx, n)
    hnd.gemm("T", "N", n, n, n, 1.0, x, n, hx, n, 0.0, hp, n)
    st = hnd.synchronize()
    if st[0]:
        raise RuntimeError(f"{spec}: S is not positive definite (leading minor {st[0]})")
    a = hp.cpu().numpy().T
    return 0.5 * (a + a.T)


def accuracy(a, wref, w, z):
    dw = float(np.max(np.abs(w - wref)))
    scale = float(np.max(np.abs(wref)))
    m = z.shape[1]
    res = float(np.max(np.abs(a @ z - z * w[:m]))) / scale
    orth = float(np.max(np.abs(z.T @ z - np.eye(m))))
    return {"max_dw": dw, "residual": res, "orth": orth}


class Timer:
    def __init__(self, torch, stream):
        self.torch, self.stream, self.marks = torch, stream, []

    def mark(self, name):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record(self.stream)
        self.marks.append((name, e))

    def stages(self):
        return {self.marks[k + 1][0]: self.marks[k][1].elapsed_time(self.marks[k + 1][1])
                for k in range(len(self.marks) - 1)}


def run_variant(torch, xtbeig, hnd, kind, prec, b, a, m):
    """One run: returns (stage ms, w (FP64 numpy), Z (n x m FP64 numpy), (d, e) bytes hash or None). Every torch
    operation runs on the library's stream, so that conversions and copies are ordered with the library's kernels."""
    stream = torch.cuda.ExternalStream(hnd.stream_ptr())
    with torch.cuda.stream(stream):
        return _run_variant(torch, xtbeig, hnd, kind, prec, b, a, m, stream)


def _run_variant(torch, xtbeig, hnd, kind, prec, b, a, m, stream):
    import ctypes
    n = a.shape[0]
    dt = torch.float64 if prec == "f64" else torch.float32
    L = xtbeig.lib()
    A = torch.from_numpy(np.ascontiguousarray(a.T)).cuda().to(dt)
    torch.cuda.synchronize()
    t = Timer(torch, stream)
    dhash = None
    t.mark("start")
    if kind == "syevd":
        w = torch.zeros(n, dtype=torch.float64, device="cuda")
        Z = torch.zeros((n, n), dtype=dt, device="cuda")
        (hnd.dsyevd if prec == "f64" else hnd.ssyevd)(A, w, Z)
        t.mark("syevd")
    else:
        d = torch.zeros(n, dtype=dt, device="cuda")
        e = torch.zeros(n, dtype=dt, device="cuda")
        if kind == "1stage":
            tau = torch.zeros(n, dtype=dt, device="cuda")
            f = L.xtbeig_dev_dsytrd_upper if prec == "f64" else L.xtbeig_dev_ssytrd_upper
            rc = f(hnd._h, n, ctypes.c_void_p(A.data_ptr()), n, ctypes.c_void_p(d.data_ptr()),
                   ctypes.c_void_p(e.data_ptr()), ctypes.c_void_p(tau.data_ptr()), 0)
            if rc:
                raise RuntimeError(xtbeig.last_error())
            t.mark("reduction")
        else:
            V1 = torch.zeros((n, n), dtype=dt, device="cuda")
            T1 = torch.zeros((n, b), dtype=dt, device="cuda")
            AB = torch.zeros(xtbeig.band_elems(n, b), dtype=dt, device="cuda")
            V2 = torch.zeros((n, n), dtype=dt, device="cuda")
            hnd.reduce_to_band(A, b, V1, T1, AB)
            t.mark("dense_to_band")
            hnd.chase(n, b, AB, V2, d, e)
            t.mark("chase")
        d64 = d.to(torch.float64, copy=True)  # the divide and conquer overwrites its inputs; d, e are hashed below
        e64 = e.to(torch.float64, copy=True)
        Zt = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        hnd.dstedc(d64, e64, Zt)
        t.mark("stedc")
        Z = Zt.to(dt)
        if kind == "1stage":
            f = L.xtbeig_dev_dormtr_upper if prec == "f64" else L.xtbeig_dev_sormtr_upper
            rc = f(hnd._h, n, m, ctypes.c_void_p(A.data_ptr()), n, ctypes.c_void_p(tau.data_ptr()),
                   ctypes.c_void_p(Z.data_ptr()), n)
            if rc:
                raise RuntimeError(xtbeig.last_error())
            t.mark("back_transform")
        else:
            hnd.back_bulges(n, b, V2, Z, m)
            t.mark("back_bulges")
            hnd.back_band(n, b, V1, T1, Z, m)
            t.mark("back_band")
        w = d64
    st = hnd.synchronize()
    if st != [0, 0, 0, 0]:
        raise RuntimeError(f"{kind}-{prec}: status {st}")
    if kind != "syevd":
        dn = d.cpu().numpy()
        en = e.cpu().numpy()[: max(n - 1, 0)]
        dhash = xtbeig.fnv1a(dn, en)
    wn = w.cpu().numpy()
    zn = Z.to(torch.float64).cpu().numpy().T[:, :m].copy()
    return t.stages(), wn, zn, dhash


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("inputs", nargs="*", default=["random:412"])
    ap.add_argument("--fraction", type=float, default=0.5)
    ap.add_argument("--bands", default="8,16,32")
    ap.add_argument("--precisions", default="f64,f32")
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)
    import scipy.linalg as sl
    import torch
    import xtbeig
    hnd = xtbeig.Handle(0)
    rows = []
    try:
        for spec in args.inputs:
            name, a = load(torch, xtbeig, hnd, spec)
            n = a.shape[0]
            m = max(1, int(round(args.fraction * n)))
            wref = sl.eigh(a, eigvals_only=True)
            variants = []
            for prec in args.precisions.split(","):
                variants.append(("1stage", prec, 0))
                for b in (int(x) for x in args.bands.split(",")):
                    if b < n - 1:
                        variants.append(("2stage", prec, b))
                variants.append(("syevd", prec, 0))
            for kind, prec, b in variants:
                label = f"{kind}-{prec}" + (f"-b{b}" if kind == "2stage" else "")
                runs = [run_variant(torch, xtbeig, hnd, kind, prec, b, a, m) for _ in range(args.reps + 1)]
                timed = runs[1:]
                stages = {k: float(np.median([r[0][k] for r in timed])) for k in timed[0][0]}
                _, w, z, dhash = runs[-1]
                row = {"input": name, "n": n, "m": m, "variant": label, "ms": stages,
                       "total_ms": float(sum(stages.values())), **accuracy(a, wref, w, z)}
                if dhash is not None:
                    row["tridiagonal_fnv1a"] = f"{dhash:016x}"
                    row["hash_stable"] = all(r[3] == dhash for r in runs)
                rows.append(row)
                print(json.dumps(row), flush=True)
    finally:
        hnd.close()
    return rows


if __name__ == "__main__":
    main()
