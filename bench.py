"""Benchmark: generalized eigensolves per second at n = 1204 on one MI355X, FP32 + FP64 refinement vs plain FP64.

One step = one generalized solve (H C = S C Lambda, all n eigenpairs) of the n = 1204 workload (H with the mixed
spectrum built from n Householder reflectors, S = I/2 + G/2 from the Gaussian overlap of points in a cube), inputs
resident in HBM. The FP64 path and the FP32+refine path are each run W untimed + K timed steps on the library's
stream, both checked against the host-LAPACK FP64 oracle (eigenvalues within 1e-12 ||H||_2). `value` is the
FP32+refine rate; the FP64 rate and their ratio are reported beside it.

Throughput beside it (not `value`): 32 solves of the same workload in flight through the stream-batched call
(throughput_concurrent) and through the grouped call, every kernel launched once for all 32 (throughput_grouped).

roofline: the dominant kernel of the FP32+refine solve (largest share of its time in the kernel trace) timed alone
with HIP events on the handle's stream, against its bound (HBM bytes or MFMA flops per launch).
cpu_baseline: the oracle (C restatement of the reference's 1-stage reduction and gemm back-transformation, with the
host LAPACK's dpotrf / dsygst / dstedc / dtrsm) on the host, one core, a bounded sample of the same workload;
cpu_baselines adds plain host dsygvd at one BLAS thread and at every host core this process may use.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "xtb-diagonalization_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

N = 1204
SEED = 1001  # configuration 1
MI355X_FP64_TFLOPS = 78.6
MI355X_FP32_TFLOPS = 157.3
MI355X_HBM_GBS = 8000.0
# HBM bytes per chase launch (FP32, n = 1204, b = 8) from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on this command,
# each scaled by the same counter's reading for a 256 MiB copy (--pmc-calibration): fetch 41.875 KiB x 2.00, write
# 2939.94 KiB x 1.00.
TRAFFIC_BYTES = 3096256


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_solves(torch, hnd, hd, sd, w, c, prec, steps, warmup):
    stream = torch.cuda.ExternalStream(hnd.stream_ptr())
    for _ in range(warmup):
        hnd.sygvd(hd, sd, w, c, prec)
    st = hnd.synchronize()
    if st != [0, 0, 0, 0]:
        raise RuntimeError(f"solver status {st}")
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    wall0 = time.perf_counter()
    for _ in range(steps):
        hnd.sygvd(hd, sd, w, c, prec)
    t1.record(stream)
    st = hnd.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    if st != [0, 0, 0, 0]:
        raise RuntimeError(f"solver status {st}")
    return t0.elapsed_time(t1) / steps, wall / steps * 1e3


def concurrent_throughput(torch, np, workloads, xtbeig, hnd, h, s, wref, count=32, slots=16):
    """`count` solves of the workload, all in flight at once through xtbeig_dev_sygvd_batched (`slots` sub-handle
    streams, HIP's default hardware queues), one warm-up pass then two timed, the rate of the faster; the first and
    last solve of the timed pass checked against the oracle. Returns {precision: {...}}."""
    n = h.shape[0]
    hd = torch.from_numpy(np.ascontiguousarray(h.T)).cuda()
    sd = torch.from_numpy(np.ascontiguousarray(s.T)).cuda()
    ws = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(count)]
    cs = [torch.zeros((n, n), dtype=torch.float64, device="cuda") for _ in range(count)]
    out = {}
    for name, prec in (("fp64", xtbeig.FP64), ("fp32_refine", xtbeig.FP32_REFINE)):
        hnd.sygvd_batched([hd] * count, [sd] * count, ws, cs, prec, slots=slots)
        hnd.synchronize()
        best = None
        for _ in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            hnd.sygvd_batched([hd] * count, [sd] * count, ws, cs, prec, slots=slots)
            st = hnd.synchronize()
            dt = time.perf_counter() - t0
            if st != [0, 0, 0, 0]:
                raise RuntimeError(f"batched status {st}")
            best = dt if best is None else min(best, dt)
        ok = all(workloads.passes(workloads.gates(h, s, ws[q].cpu().numpy(), cs[q].cpu().numpy().T, wref))
                 for q in (0, count - 1))
        out[name] = {"solves_per_s": count / best, "wall_ms": best * 1e3, "checked_pass": ok}
        log(f"throughput {name}", out[name])
    out["config"] = {"solves_in_flight": count, "slots": slots, "api": "xtbeig_dev_sygvd_batched",
                     "timing": "host wall clock around the call, inputs resident in HBM"}
    return out


def grouped_throughput(torch, np, workloads, xtbeig, hnd, h, s, wref, count=32):
    """`count` solves of the workload through xtbeig_dev_sygvd_grouped (no padding: one group, every kernel of the
    solve launched once for all of them), one warm-up pass then two timed, the rate of the faster; every solve's status
    words checked and the first and last solve gated against the oracle. Returns {precision: {...}}."""
    n = h.shape[0]
    hd = torch.from_numpy(np.ascontiguousarray(h.T)).cuda()
    sd = torch.from_numpy(np.ascontiguousarray(s.T)).cuda()
    ws = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(count)]
    cs = [torch.zeros((n, n), dtype=torch.float64, device="cuda") for _ in range(count)]
    out = {}
    for name, prec in (("fp64", xtbeig.FP64), ("fp32_refine", xtbeig.FP32_REFINE)):
        best = None
        for rep in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st, _ = hnd.sygvd_grouped([hd] * count, [sd] * count, ws, cs, prec, granule=1)
            dt = time.perf_counter() - t0
            if any(x != [0, 0, 0, 0] for x in st):
                raise RuntimeError(f"grouped status {st}")
            if rep:
                best = dt if best is None else min(best, dt)
        ok = all(workloads.passes(workloads.gates(h, s, ws[q].cpu().numpy(), cs[q].cpu().numpy().T, wref))
                 for q in (0, count - 1))
        out[name] = {"solves_per_s": count / best, "wall_ms": best * 1e3, "checked_pass": ok}
        log(f"grouped throughput {name}", out[name])
    out["config"] = {"solves_per_call": count, "api": "xtbeig_dev_sygvd_grouped", "granule": 1,
                     "timing": "host wall clock around the call, inputs resident in HBM"}
    return out


def kernel_roofline(torch, xtbeig, hnd, h, s, b):
    """The bulge chase of the FP32+refine solve (one workgroup, band resident in LDS) is its largest kernel. It is
    timed alone, with HIP events on the handle's stream, on the FP32 band that the library's own stage 1 makes from
    H' = X^T H X. Algorithmic bytes per launch: the chase streams 2 (2b + 1) n^2 elements (read + write of the band
    window of every bulge), the count charged to it when nothing stays in LDS or cache."""
    import numpy as np
    n = N
    stream = torch.cuda.ExternalStream(hnd.stream_ptr())
    # torch's own operations below run on the library's stream, ordered with its kernels
    ctx = torch.cuda.stream(stream)
    ctx.__enter__()
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
    a32 = hp.float()
    v1 = torch.zeros((n, n), dtype=torch.float32, device="cuda")
    t1 = torch.zeros((n, b), dtype=torch.float32, device="cuda")
    ab = torch.zeros(xtbeig.band_elems(n, b), dtype=torch.float32, device="cuda")
    hnd.reduce_to_band(a32, b, v1, t1, ab)
    hnd.synchronize()
    v2 = torch.zeros((n, n), dtype=torch.float32, device="cuda")
    d = torch.zeros(n, dtype=torch.float32, device="cuda")
    e = torch.zeros(n, dtype=torch.float32, device="cuda")
    reps = 5
    abw = [ab.clone() for _ in range(reps + 1)]
    hnd.synchronize()
    torch.cuda.synchronize()
    ctx.__exit__(None, None, None)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * (reps + 1))]
    for r in range(reps + 1):
        ev[2 * r].record(stream)
        hnd.chase(n, b, abw[r], v2, d, e)
        ev[2 * r + 1].record(stream)
    st = hnd.synchronize()
    torch.cuda.synchronize()
    if st != [0, 0, 0, 0]:
        raise RuntimeError(f"chase status {st}")
    times = [ev[2 * r].elapsed_time(ev[2 * r + 1]) for r in range(1, reps + 1)]
    ms = sum(times) / len(times)
    bytes_per_launch = 2 * (2 * b + 1) * n * n * 4
    achieved = bytes_per_launch / (ms * 1e-3) / 1e9
    return {"kernel": f"chase_lanes_kernel<float, {b}, 2, LDS> (bulge chase, n={n}, b={b})", "bound": "hbm",
            "achieved": achieved, "peak": MI355X_HBM_GBS, "unit": "GB/s", "frac": achieved / MI355X_HBM_GBS,
            "traffic": TRAFFIC_BYTES, "ms_per_launch": ms, "bytes_per_launch": bytes_per_launch}


def pmc_calibration(torch):
    """A plain 256 MiB device copy (one read, one write of 256 MiB), for scaling FETCH_SIZE / WRITE_SIZE counters."""
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()


def cpu_baseline(h, s, sample=5):
    """The oracle's FP64 generalized solve on one host core: host dpotrf + dsygst, the C restatement of sytrd.c's
    reduction (LAPACK dsytrd 'U', nb = 32), host dstedc, the restated gemm back-transformation, host dtrsm."""
    import numpy as np
    import host_lapack as hl
    import sbr
    hl.set_threads(1)
    times = []
    w = None
    for _ in range(sample):
        t0 = time.perf_counter()
        u, info = hl.dpotrf_upper(s)
        u = np.triu(u)
        hp, info = hl.dsygst_upper(h, u)
        a, d, e, tau = sbr.sytrd_upper(hp)
        w, zt, info = hl.dstedc(d, e)
        z = sbr.ormtr_upper(a, tau, zt)
        c = hl.dtrsm_left_upper(u, z)
        times.append(time.perf_counter() - t0)
    return sum(times) / len(times), w


def cpu_lapack_baseline(h, s, threads, sample=5):
    """Plain host LAPACK dsygvd(1, 'V', 'U') on `threads` BLAS threads: the median of `sample` solves after one
    untimed warm-up."""
    import numpy as np
    import host_lapack as hl
    hl.set_threads(threads)
    try:
        w, _, info = hl.dsygvd(h, s)
        if info != 0:
            raise RuntimeError(f"host dsygvd info {info}")
        times = []
        for _ in range(sample):
            t0 = time.perf_counter()
            w, _, info = hl.dsygvd(h, s)
            times.append(time.perf_counter() - t0)
    finally:
        hl.set_threads(1)
    return float(np.median(times)), w


def host_cores():
    """CPU threads this process may use: the affinity mask, capped by OMP_NUM_THREADS when that is set."""
    cores = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        cores = min(cores, int(omp))
    return cores


def kernels(torch, np, workloads, xtbeig, hnd, h, s, which, reps=5):
    """The hot kernels at n = 1204, each run `reps` times with nothing else in flight, for counter collection
    (rocprofv3 --pmc): FP64 and FP32 GEMM n^3 (gemm_kernel), the dense-to-band reduction at b = 8 in both
    precisions (panel kernels), the FP32 bulge chase at b = 8, the FP64 and FP32+refine generalized solves of the
    bench workload (gemm_upper_kernel, the refinement and check kernels, every stage in place), and the one-stage
    reduction (trd_symv_kernel and the dlatrd column kernels). `which`: the groups to run (gemm, band, chase, solve,
    sytrd). Prints one line per stage."""
    n = N
    for dt in ((torch.float64, torch.float32) if "gemm" in which else ()):
        a = torch.randn(n, n, dtype=dt, device="cuda")
        b = torch.randn(n, n, dtype=dt, device="cuda")
        c = torch.zeros(n, n, dtype=dt, device="cuda")
        for _ in range(reps):
            hnd.gemm("N", "N", n, n, n, 1.0, a, n, b, n, 0.0, c, n)
        hnd.synchronize()
        log(f"gemm NN {dt} x{reps}")
    a, _ = workloads.standard_matrix(n, 11)
    for dt in ((torch.float64, torch.float32) if ("band" in which or "chase" in which) else ()):
        bw = xtbeig.default_band(n, fp32=dt == torch.float32)
        A0 = torch.from_numpy(np.ascontiguousarray(a.T)).cuda().to(dt)
        V1 = torch.zeros((n, n), dtype=dt, device="cuda")
        T1 = torch.zeros((n, bw), dtype=dt, device="cuda")
        AB = torch.zeros(xtbeig.band_elems(n, bw), dtype=dt, device="cuda")
        for _ in range(reps):
            A = A0.clone()
            torch.cuda.synchronize()
            hnd.reduce_to_band(A, bw, V1, T1, AB)
            hnd.synchronize()
        log(f"reduce_to_band b={bw} {dt} x{reps}")
        if dt == torch.float32 and "chase" in which:
            AB0 = AB.clone()
            V2 = torch.zeros((n, n), dtype=dt, device="cuda")
            D = torch.zeros(n, dtype=dt, device="cuda")
            E = torch.zeros(n, dtype=dt, device="cuda")
            for _ in range(reps):
                ABc = AB0.clone()
                torch.cuda.synchronize()
                hnd.chase(n, bw, ABc, V2, D, E)
                hnd.synchronize()
            log(f"chase b={bw} {dt} x{reps}")
    if "solve" in which:
        hd = torch.from_numpy(np.ascontiguousarray(h.T)).cuda()
        sd = torch.from_numpy(np.ascontiguousarray(s.T)).cuda()
        w = torch.zeros(n, dtype=torch.float64, device="cuda")
        c = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        for prec in (xtbeig.FP64, xtbeig.FP32_REFINE):
            for _ in range(reps):
                hnd.sygvd(hd, sd, w, c, prec)
                st = hnd.synchronize()
                if st != [0, 0, 0, 0]:
                    raise RuntimeError(f"sygvd status {st}")
            log(f"sygvd precision {prec} x{reps}")
    if "sytrd" not in which:
        return
    import ctypes
    L = xtbeig.lib()
    A0 = torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
    d = torch.zeros(n, dtype=torch.float64, device="cuda")
    e = torch.zeros(n, dtype=torch.float64, device="cuda")
    tau = torch.zeros(n, dtype=torch.float64, device="cuda")
    for _ in range(2):
        A = A0.clone()
        torch.cuda.synchronize()
        rc = L.xtbeig_dev_dsytrd_upper(hnd._h, n, ctypes.c_void_p(A.data_ptr()), n, ctypes.c_void_p(d.data_ptr()),
                                       ctypes.c_void_p(e.data_ptr()), ctypes.c_void_p(tau.data_ptr()), 0)
        if rc != 0:
            raise RuntimeError(xtbeig.last_error())
        hnd.synchronize()
    log("one-stage dsytrd x2")


def band_sweep(torch, np, workloads, xtbeig, hnd, sizes=(1204, 1788), bands=(8, 16, 32, 48, 64)):
    """The generalized solve (FP64 and FP32+refine, workload pair of seed 1001 + n) and its band-dependent stages
    alone (dense -> band, chase, bulge back-transform of m = n columns, each precision; standard matrix of seed 11)
    for each band width, HIP events on the handle's stream, median of 3 after a warm-up. One JSON line per (n, b)."""
    stream = torch.cuda.ExternalStream(hnd.stream_ptr())

    def ev_time(fn, reps=3):
        ts = []
        for r in range(reps + 1):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            st = hnd.synchronize()
            if st != [0, 0, 0, 0]:
                raise RuntimeError(f"status {st}")
            if r:
                ts.append(e0.elapsed_time(e1))
        return float(np.median(ts))

    with torch.cuda.stream(stream):
        for n in sizes:
            h, s = workloads.generalized_pair(n, 1001 + n)
            hd = torch.from_numpy(np.ascontiguousarray(h.T)).cuda()
            sd = torch.from_numpy(np.ascontiguousarray(s.T)).cuda()
            w = torch.zeros(n, dtype=torch.float64, device="cuda")
            c = torch.zeros((n, n), dtype=torch.float64, device="cuda")
            a, _ = workloads.standard_matrix(n, 11)
            for b in bands:
                row = {"n": n, "band": b}
                hnd.set_tuning(b, 2)
                for name, prec in (("sygvd_fp64_ms", xtbeig.FP64), ("sygvd_fp32_refine_ms", xtbeig.FP32_REFINE)):
                    row[name] = round(ev_time(lambda: hnd.sygvd(hd, sd, w, c, prec)), 2)
                for dt, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
                    A0 = torch.from_numpy(np.ascontiguousarray(a.T)).cuda().to(dt)
                    A = A0.clone()
                    V1 = torch.zeros((n, n), dtype=dt, device="cuda")
                    T1 = torch.zeros((n, b), dtype=dt, device="cuda")
                    AB = torch.zeros(xtbeig.band_elems(n, b), dtype=dt, device="cuda")
                    row[f"dense_to_band_{tag}_ms"] = round(
                        ev_time(lambda: (A.copy_(A0), hnd.reduce_to_band(A, b, V1, T1, AB))), 2)
                    AB0 = AB.clone()
                    V2 = torch.zeros((n, n), dtype=dt, device="cuda")
                    D = torch.zeros(n, dtype=dt, device="cuda")
                    E = torch.zeros(n, dtype=dt, device="cuda")
                    row[f"chase_{tag}_ms"] = round(ev_time(lambda: (AB.copy_(AB0), hnd.chase(n, b, AB, V2, D, E))), 2)
                    row[f"chase_{tag}_band_in_lds"] = bool(xtbeig.lib().xtbeig_chase_in_lds(n, b, int(dt == torch.float32)))
                    Z = torch.eye(n, dtype=dt, device="cuda")
                    row[f"bulge_back_{tag}_ms"] = round(ev_time(lambda: hnd.back_bulges(n, b, V2, Z, n)), 2)
                print(json.dumps(row), flush=True)
            hnd.set_tuning(0, 2)


def rocprof_summary(path, top=40):
    """Markdown tables of a rocprofv3 --output-format csv directory: kernel statistics (from --stats) and, for a
    --pmc run, per kernel and grid size the dispatch count, mean duration and mean of each counter, plus MFMA rates (MOPS x 512
    flops; MFMA busy cycles over 256 CUs x 4 SIMDs x the per-XCD GRBM_GUI_ACTIVE) and the largest single
    __amd_rocclr_copyBuffer / fill dispatch (the --pmc-calibration copy of 256 MiB) for scaling FETCH_SIZE and
    WRITE_SIZE. Host-side post-processing only."""
    import collections
    import csv
    import glob
    out = []
    for f in glob.glob(os.path.join(path, "**", "*kernel_stats.csv"), recursive=True):
        rows = list(csv.DictReader(open(f)))
        tot = sum(float(r["TotalDurationNs"]) for r in rows)
        out.append(f"Kernel time summed over all dispatches: {tot / 1e6:.2f} ms.\n")
        out.append("| kernel | calls | total us | avg us | % |\n|---|---:|---:|---:|---:|")
        for r in rows[:top]:
            out.append(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e3:.1f} | "
                       f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
    for f in glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True):
        per = collections.defaultdict(dict)  # (kernel, grid, dispatch) -> counter -> value
        dur = {}
        for r in csv.DictReader(open(f)):
            key = (_short(r["Kernel_Name"]), r.get("Grid_Size", "?"), r["Dispatch_Id"])
            per[key][r["Counter_Name"]] = per[key].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
            if r.get("Start_Timestamp"):
                dur[key] = float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
        names = sorted({c for v in per.values() for c in v})
        # the GEMM-shaped kernels together (every product of the solve: plain, upper-tile, split-K, thin and the
        # D&C merge products): MFMA flops over their summed durations
        # (each precision over the kernels that run its MFMA instructions)
        parts = []
        for prec, peak in (("F64", MI355X_FP64_TFLOPS), ("F32", MI355X_FP32_TFLOPS)):
            ns, fl = 0.0, 0.0
            for key, v in per.items():
                f = v.get(f"SQ_INSTS_VALU_MFMA_MOPS_{prec}", 0.0) * 512
                if "gemm" in key[0] and dur.get(key) and f > 0:
                    ns += dur[key]
                    fl += f
            if ns > 0:
                parts.append(f"{prec}: {fl / 1e9:.1f} GFLOP in {ns / 1e6:.2f} ms, {fl / ns / 1e3:.1f} TF/s = "
                             f"{fl / ns / 1e3 / peak:.1%} of peak")
        if parts:
            out.append("\nGEMM-shaped kernels together: " + "; ".join(parts))
        for by_grid, title, rows in ((False, "per kernel", top), (True, "per kernel and grid size", top // 2)):
            agg = collections.defaultdict(list)
            for key, v in per.items():
                agg[f"{key[0]} [grid {key[1]}]" if by_grid else key[0]].append((v, dur.get(key)))
            out.append(f"\nCounters {title} (means over dispatches):\n")
            out.append("| kernel | dispatches | mean us | " + " | ".join(names) + " | derived |")
            out.append("|---|---:|---:|" + "---:|" * len(names) + "---|")
            order = sorted(agg, key=lambda k: -sum((d or 0) for _, d in agg[k]))
            for k in order[:rows]:
                out.append(_counter_row(k, agg[k], names))
    return "\n".join(out)


def _counter_row(k, vs, names):
    """One markdown row: means of each counter over the dispatches vs, and the rates derived from them."""
    ds = [d for _, d in vs if d]
    mean_ns = sum(ds) / len(ds) if ds else float("nan")
    means = {c: sum(v.get(c, 0.0) for v, _ in vs) / len(vs) for c in names}
    derived = []
    for prec, peak in (("F64", MI355X_FP64_TFLOPS), ("F32", MI355X_FP32_TFLOPS)):
        mops = means.get(f"SQ_INSTS_VALU_MFMA_MOPS_{prec}", 0.0)
        if mops > 0 and ds:
            tf = mops * 512 / mean_ns / 1e3
            derived.append(f"{prec}: {mops * 512 / 1e9:.3f} GFLOP, {tf:.1f} TF/s = {tf / peak:.1%} of peak")
    if means.get("SQ_VALU_MFMA_BUSY_CYCLES", 0) > 0 and means.get("GRBM_GUI_ACTIVE", 0) > 0:
        util = means["SQ_VALU_MFMA_BUSY_CYCLES"] / (means["GRBM_GUI_ACTIVE"] / 8 * 1024)
        derived.append(f"MFMA busy {util:.1%}")
    if "SQ_INSTS_LDS_LOAD_BANDWIDTH" in means and ds:
        lds_b = 64 * (means["SQ_INSTS_LDS_LOAD_BANDWIDTH"] + means.get("SQ_INSTS_LDS_STORE_BANDWIDTH", 0))
        derived.append(f"LDS {lds_b / 1e6:.1f} MB, {lds_b / mean_ns:.1f} GB/s")
    if ("copyBuffer" in k or "Fill" in k) and any(c in names for c in ("FETCH_SIZE", "WRITE_SIZE")):
        big = max(vs, key=lambda t: t[0].get("FETCH_SIZE", 0.0) + t[0].get("WRITE_SIZE", 0.0))[0]
        derived.append("largest dispatch: " + ", ".join(f"{c} {big[c]:.1f} KiB" for c in
                                                        ("FETCH_SIZE", "WRITE_SIZE") if c in big))
    return (f"| `{k}` | {len(vs)} | {mean_ns / 1e3:.2f} | " + " | ".join(f"{means[c]:.1f}" for c in names) + " | " +
            "; ".join(derived) + " |")


def _short(name):
    """A kernel's name without its namespace and argument list."""
    name = name.replace("void ", "").replace("xtbeig::(anonymous namespace)::", "")
    depth, cut = 0, len(name)
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    return name[:cut][:110]


def components(torch, np, workloads, xtbeig, hnd, h, s):
    """Each entry point of the path timed alone at n = 1204, one JSON line per entry point.

    Device-resident stages: HIP events on the handle's stream around each call, inputs restored (outside the timed
    region) before every call, the median of 7 calls after 2 warm-up ones. Host-array entry points (the sytrd.c
    symbol set and the LAPACK-style drivers): wall clock around the call, host-device transfers included.
    """
    n = N
    stream = torch.cuda.ExternalStream(hnd.stream_ptr())

    def dev_time(restore, call, reps=7, warm=2):
        ts = []
        for r in range(warm + reps):
            restore()
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            call()
            e1.record(stream)
            st = hnd.synchronize()
            if st != [0, 0, 0, 0]:
                raise RuntimeError(f"status {st}")
            if r >= warm:
                ts.append(e0.elapsed_time(e1))
        return float(np.median(ts))

    def emit(row, name, ms, dtype, note):
        print(json.dumps({"component": name, "path_row": row, "n": n, "ms": round(ms, 4), "dtype": dtype,
                          "timing": note}), flush=True)

    a, _ = workloads.standard_matrix(n, 11)
    for dt, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
        b = xtbeig.default_band(n, fp32=dt == torch.float32)
        A0 = torch.from_numpy(np.ascontiguousarray(a.T)).cuda().to(dt)
        A = A0.clone()
        V1 = torch.zeros((n, n), dtype=dt, device="cuda")
        T1 = torch.zeros((n, b), dtype=dt, device="cuda")
        AB = torch.zeros(xtbeig.band_elems(n, b), dtype=dt, device="cuda")
        ms = dev_time(lambda: A.copy_(A0), lambda: hnd.reduce_to_band(A, b, V1, T1, AB))
        emit("a7", f"reduce_to_band (dense -> band {b})", ms, tag, "HIP events, device-resident")
        AB0 = AB.clone()
        V2 = torch.zeros((n, n), dtype=dt, device="cuda")
        D = torch.zeros(n, dtype=dt, device="cuda")
        E = torch.zeros(n, dtype=dt, device="cuda")
        ms = dev_time(lambda: AB.copy_(AB0), lambda: hnd.chase(n, b, AB, V2, D, E))
        emit("a8", f"chase (band {b} -> tridiagonal)", ms, tag, "HIP events, device-resident")
        Z = torch.eye(n, dtype=dt, device="cuda")
        ms = dev_time(lambda: None, lambda: hnd.back_bulges(n, b, V2, Z, n))
        emit("a9", "bulge back-transform Z <- Q2 Z (m = n)", ms, tag, "HIP events, device-resident")
        ms = dev_time(lambda: None, lambda: hnd.back_band(n, b, V1, T1, Z, n))
        emit("a10/a11", "band back-transform Z <- Q1 Z (m = n, WY blocks)", ms, tag, "HIP events, device-resident")
        W = torch.zeros(n, dtype=torch.float64, device="cuda")
        Zs = torch.zeros((n, n), dtype=dt, device="cuda")
        ms = dev_time(lambda: A.copy_(A0),
                      lambda: (hnd.dsyevd if dt == torch.float64 else hnd.ssyevd)(A, W, Zs))
        emit("a12", "syevd (standard eigensolve, two-stage chain, all eigenpairs)", ms, tag,
             "HIP events, device-resident")
    d0 = torch.from_numpy(np.linspace(-1.0, 1.0, n) + 0.01 * np.sin(np.arange(n))).cuda()
    e0 = torch.full((n,), 0.3, dtype=torch.float64, device="cuda")
    d, e = d0.clone(), e0.clone()
    Zt = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    ms = dev_time(lambda: (d.copy_(d0), e.copy_(e0)), lambda: hnd.dstedc(d, e, Zt))
    emit("a5", "stedc (divide and conquer, all eigenpairs of T)", ms, "f64", "HIP events, device-resident")
    S0 = torch.from_numpy(np.ascontiguousarray(s.T)).cuda()
    S = S0.clone()
    U = torch.zeros_like(S)
    X = torch.zeros_like(S)
    ms = dev_time(lambda: S.copy_(S0), lambda: hnd.potrf_inverse(S, U, X))
    emit("a13", "potrf + inverse of S (U, X = U^-1)", ms, "f64", "HIP events, device-resident")
    for dt, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
        Ga = torch.randn(n, n, dtype=dt, device="cuda")
        Gb = torch.randn(n, n, dtype=dt, device="cuda")
        Gc = torch.zeros(n, n, dtype=dt, device="cuda")
        ms = dev_time(lambda: None, lambda: hnd.gemm("N", "N", n, n, n, 1.0, Ga, n, Gb, n, 0.0, Gc, n))
        emit("a11/a13/a15", f"gemm NN {n}^3 ({2 * n ** 3 / ms / 1e9:.1f} TF/s)", ms, tag,
             "HIP events, device-resident")
    # host-array entry points (the sytrd.c symbol set and LAPACK-style drivers), transfers included: the C calls
    # alone, on caller arrays allocated once (inputs restored outside the timed region)
    import ctypes
    L = xtbeig.lib()
    I, P = xtbeig._i, xtbeig._ptr

    def host_call(restore, call, reps=3):
        ts = []
        for _ in range(reps + 1):
            restore()
            t0 = time.perf_counter()
            call()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts[1:])) * 1e3

    for dtype, pre in ((np.float64, "d"), (np.float32, "s")):
        tag = "f64" if pre == "d" else "f32"
        a0 = np.asfortranarray(a, dtype=dtype)
        A_ = a0.copy(order="F")
        d_, e_, tau_ = (np.zeros(n, dtype=dtype) for _ in range(3))
        work = np.zeros(32 * n, dtype=dtype)
        ms = host_call(lambda: np.copyto(A_, a0),
                       lambda: getattr(L, f"xtb_{pre}sytrd_upper")(I(n), P(A_), I(n), P(d_), P(e_), P(tau_), P(work)))
        emit("a1-a4", f"xtb_{pre}sytrd_upper (one-stage reduction)", ms, tag, "wall clock, host arrays")
        d0, e0 = d_.copy(), e_.copy()
        Zt = np.zeros((n, n), dtype=dtype, order="F")
        lw, liw = 1 + 4 * n + n * n, 3 + 5 * n
        wk, iwk, info = np.zeros(lw, dtype=dtype), np.zeros(liw, dtype=np.int32), ctypes.c_int(0)
        ms = host_call(lambda: (np.copyto(d_, d0), np.copyto(e_, e0)),
                       lambda: getattr(L, f"xtb_{pre}stedc")(I(n), P(d_), P(e_), P(Zt), I(n), P(wk), I(lw), P(iwk),
                                                            I(liw), ctypes.byref(info)))
        emit("a5", f"xtb_{pre}stedc (host arrays)", ms, tag, "wall clock, host arrays")
        Z0 = np.asfortranarray(np.eye(n, dtype=dtype))
        Z_ = Z0.copy(order="F")
        wk2 = np.zeros(32 * (4 * n + 32), dtype=dtype)
        ms = host_call(lambda: np.copyto(Z_, Z0),
                       lambda: getattr(L, f"xtb_{pre}ormtr_upper")(I(n), I(n), P(A_), I(n), P(tau_), P(Z_), I(n),
                                                                  P(wk2)))
        emit("a6", f"xtb_{pre}ormtr_upper (m = n)", ms, tag, "wall clock, host arrays")
    a0 = np.asfortranarray(a, dtype=np.float64)
    A_ = a0.copy(order="F")
    d_, e_, tau_ = (np.zeros(n) for _ in range(3))
    work = np.zeros(32 * n)
    ms = host_call(lambda: np.copyto(A_, a0),
                   lambda: L.xtb_dsytrd_full(I(n), P(A_), I(n), P(d_), P(e_), P(tau_), P(work)))
    emit("a14", "xtb_dsytrd_full (full padded copy)", ms, "f64", "wall clock, host arrays")
    Dp = np.zeros((n + 8, n), order="F")
    ms = host_call(lambda: None, lambda: L.xtb_dsymmetrize_pad(I(n), P(a0), I(n), P(Dp), I(n + 8)))
    emit("a14", "xtb_dsymmetrize_pad", ms, "f64", "wall clock, host arrays")
    lw, liw = 1 + 6 * n + 2 * n * n, 3 + 5 * n
    wk, iwk, info = np.zeros(lw), np.zeros(liw, dtype=np.int32), ctypes.c_int(0)
    w_ = np.zeros(n)
    ms = host_call(lambda: np.copyto(A_, a0),
                   lambda: L.xtbeig_dsyevd(b"V", b"U", I(n), P(A_), I(n), P(w_), P(wk), I(lw), P(iwk), I(liw),
                                           ctypes.byref(info)))
    emit("a12", "xtbeig_dsyevd (host arrays)", ms, "f64", "wall clock, host arrays")
    h0, s0 = np.asfortranarray(h), np.asfortranarray(s)
    H_, S_ = h0.copy(order="F"), s0.copy(order="F")
    for fn, row, name, dt in (("xtbeig_dsygvd", "a13", "xtbeig_dsygvd (host arrays)", "f64"),
                              ("xtbeig_dsygvd_mixed", "a13/a15", "xtbeig_dsygvd_mixed (host arrays, FP32 + refinement)",
                               "f32+f64")):
        ms = host_call(lambda: (np.copyto(H_, h0), np.copyto(S_, s0)),
                       lambda: getattr(L, fn)(I(1), b"V", b"U", I(n), P(H_), I(n), P(S_), I(n), P(w_), P(wk), I(lw),
                                              P(iwk), I(liw), ctypes.byref(info)))
        if info.value != 0:
            raise RuntimeError(f"{fn} info {info.value}")
        emit(row, name, ms, dt, "wall clock, host arrays")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="leave out the CPU baseline (for profiler runs)")
    ap.add_argument("--cpu-sample", type=int, default=5, help="solves in the CPU baseline's sample")
    ap.add_argument("--pmc-calibration", action="store_true", help="run a 256 MiB copy first (counter scaling)")
    ap.add_argument("--components", action="store_true",
                    help="time each entry point of the path alone at n = 1204 (one JSON line each) and exit")
    ap.add_argument("--kernels", default=None, metavar="GROUPS",
                    help="run the hot kernels alone a few times each (for rocprofv3 counter runs) and exit; GROUPS "
                         "is a comma-separated subset of gemm,band,chase,solve,sytrd")
    ap.add_argument("--bands", action="store_true",
                    help="time the solve and its band-dependent stages for band widths 8 to 64 at n = 1204 and 1788 "
                         "(one JSON line each) and exit")
    ap.add_argument("--rocprof-summary", default=None, metavar="DIR",
                    help="print markdown tables of a rocprofv3 csv output directory and exit (no GPU use)")
    args = ap.parse_args()
    if args.rocprof_summary:
        print(rocprof_summary(args.rocprof_summary))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the eigensolves are "
              "independent per GPU and would shard by problem with no exchange.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import scipy.linalg as sl
    import torch
    import workloads
    import xtbeig

    log("generating workload n=1204 ...")
    h, s = workloads.generalized_pair(N, SEED)
    nh = np.linalg.norm(h, 2)
    wref = sl.eigh(h, s, eigvals_only=True)
    hd = torch.from_numpy(np.ascontiguousarray(h.T)).cuda()
    sd = torch.from_numpy(np.ascontiguousarray(s.T)).cuda()
    w = torch.zeros(N, dtype=torch.float64, device="cuda")
    c = torch.zeros((N, N), dtype=torch.float64, device="cuda")
    hnd = xtbeig.Handle(0)
    if args.pmc_calibration:
        pmc_calibration(torch)
    if args.components:
        components(torch, np, workloads, xtbeig, hnd, h, s)
        return
    if args.kernels:
        kernels(torch, np, workloads, xtbeig, hnd, h, s, set(args.kernels.split(",")))
        return
    if args.bands:
        band_sweep(torch, np, workloads, xtbeig, hnd)
        return
    band = xtbeig.default_band(N, fp32=True)

    res = {}
    for name, prec in (("fp64", xtbeig.FP64), ("fp32_refine", xtbeig.FP32_REFINE)):
        log(f"timing {name} ...")
        ms, wall_ms = timed_solves(torch, hnd, hd, sd, w, c, prec, args.steps, args.warmup)
        g = workloads.gates(h, s, w.cpu().numpy(), c.cpu().numpy().T, wref)
        res[name] = {"ms": ms, "wall_ms": wall_ms, "solves_per_s": 1e3 / ms, "gates": g,
                     "pass": workloads.passes(g)}
        log(name, res[name])

    # the FP32 solve's error against the oracle before refinement (untimed): the same call with no refinement step
    hnd.set_tuning(0, 0)
    hnd.sygvd(hd, sd, w, c, xtbeig.FP32_REFINE)
    hnd.synchronize()
    unrefined = workloads.gates(h, s, w.cpu().numpy(), c.cpu().numpy().T, wref)
    hnd.set_tuning(0, 2)
    log("fp32 before refinement", unrefined)

    log("concurrent throughput ...")
    thru = concurrent_throughput(torch, np, workloads, xtbeig, hnd, h, s, wref)
    thru["ratio_fp32_refine_vs_fp64"] = thru["fp32_refine"]["solves_per_s"] / thru["fp64"]["solves_per_s"]
    log("grouped throughput ...")
    grouped = grouped_throughput(torch, np, workloads, xtbeig, hnd, h, s, wref)
    grouped["ratio_fp32_refine_vs_fp64"] = grouped["fp32_refine"]["solves_per_s"] / grouped["fp64"]["solves_per_s"]

    log("kernel roofline ...")
    roof = kernel_roofline(torch, xtbeig, hnd, h, s, band)
    log(roof)

    cpu = None
    cpu_all = None
    if not args.no_cpu:
        import host_lapack as hl
        log("cpu baseline (oracle, 1 core) ...")
        t, wc = cpu_baseline(h, s, sample=args.cpu_sample)
        cpu = {"value": 1.0 / t, "unit": "gen. eigensolves/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_sample} FP64 generalized solves at n=1204 of the same workload (oracle: host "
                         "dpotrf/dsygst/dstedc/dtrsm, one thread, + C restatement of the 1-stage reduction and "
                         "gemm back-transform), rate of the mean",
               "seconds_per_solve": t,
               "eig_err_vs_lapack": float(np.max(np.abs(wc - wref)) / nh)}
        log(cpu)
        cores = host_cores()
        cpu_all = [cpu]
        for th in sorted({1, cores}):
            log(f"cpu baseline (host dsygvd, {th} threads) ...")
            tl, wl = cpu_lapack_baseline(h, s, th, sample=args.cpu_sample)
            cpu_all.append({"value": 1.0 / tl, "unit": "gen. eigensolves/s", "cores": th, "kind": "host_lapack",
                            "sample": f"{args.cpu_sample} plain host dsygvd(1, 'V', 'U') solves at n=1204 of the same "
                                      "workload after one warm-up, rate of the median",
                            "seconds_per_solve": tl, "blas_threads": th, "library": os.path.basename(hl.LIBRARY),
                            "host_cpus_affinity": len(os.sched_getaffinity(0)), "nproc": os.cpu_count(),
                            "eig_err_vs_lapack": float(np.max(np.abs(wl - wref)) / nh)})
            log(cpu_all[-1])

    mixed, fp64 = res["fp32_refine"], res["fp64"]
    flops = 7.0 * N ** 3
    out = {
        "metric": "gen. eigensolves/s n=1204: FP32+refine vs FP64 GPU at eig err <=1e-12*||H||",
        "value": mixed["solves_per_s"],
        "unit": "gen. eigensolves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "generalized symmetric-definite eigensolve (sygvd, all eigenpairs), n=1204, H with "
                               "the mixed spectrum from n Householder reflectors, S=I/2+G/2 Gaussian overlap, seed 1001",
                   "n": N, "precision": "FP32 reduction/back-transforms + FP64 Cholesky, D&C, OA refinement"},
        "fp64": {"solves_per_s": fp64["solves_per_s"], "ms_per_step": fp64["ms"], "pass": fp64["pass"],
                 "gates": fp64["gates"]},
        "fp32_refine": {"solves_per_s": mixed["solves_per_s"], "ms_per_step": mixed["ms"], "pass": mixed["pass"],
                        "gates": mixed["gates"]},
        "fp32_before_refinement": {"gates": unrefined, "pass": workloads.passes(unrefined)},
        "speedup_fp32_refine_vs_fp64": fp64["ms"] / mixed["ms"],
        "throughput_concurrent": thru,
        "throughput_grouped": grouped,
        "nominal_flop_rate_TFLOPs": {"fp64": flops / (fp64["ms"] * 1e-3) / 1e12,
                                     "fp32_refine": flops / (mixed["ms"] * 1e-3) / 1e12},
        "roofline": {k: roof[k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "roofline_kernel": roof["kernel"],
        "roofline_ms_per_launch": roof["ms_per_launch"],
        "cpu_baseline": cpu,
        "cpu_baselines": cpu_all,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
