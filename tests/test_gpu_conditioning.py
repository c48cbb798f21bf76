"""Generalized solves with the overlap conditioning the configurations state (cond_2(S) = 1e3 and 1e4), the
refinement's cluster step and a-posteriori check, and the FP64 fallbacks the check triggers.

Pins: the FP64 host LAPACK (scipy.linalg.eigh) with the contract's gates; for overlap-weighted H the generalized
spectrum is also known by construction (the configuration spectrum, with exactly repeated pairs and pairs split by
1e-9). The refinement check's ratios are compared with the gates computed on the host.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sl

import workloads
import xtbeig

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda()


def solve(torch, handle, h, s, prec):
    n = h.shape[0]
    w = torch.zeros(n, dtype=torch.float64, device="cuda")
    c = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    hd, sd = dev(torch, h), dev(torch, s)  # referenced until the solve is done (it runs on the library's stream)
    handle.sygvd(hd, sd, w, c, prec)
    st = handle.synchronize()
    return w.cpu().numpy(), c.cpu().numpy().T, st


def cond2(s):
    e = np.linalg.eigvalsh(s)
    return e[-1] / e[0]


@pytest.mark.parametrize("n", [1204, 1788])
@pytest.mark.parametrize("cond", [1e3, 1e4])
def test_overlap_weighted_ill_conditioned(handle, torch_cuda, n, cond):
    """H = U^T M U with cond_2(S) = cond: both precisions pass the gates against host eigh, the eigenvalues match the
    construction, and the refinement check agrees with the host gates."""
    torch = torch_cuda
    handle.set_tuning(0, 2)
    h, s, lam = workloads.generalized_pair_cond(n, workloads.SEED_BASE + 1, cond, physical=True)
    c2 = cond2(s)
    print(f"n={n} cond2(S)={c2:.4e}", flush=True)
    assert abs(c2 / cond - 1.0) < 1e-6
    wref = sl.eigh(h, s, eigvals_only=True)
    assert np.max(np.abs(wref - lam)) < 1e-12 * np.linalg.norm(h, 2)
    for prec in (xtbeig.FP64, xtbeig.FP32_REFINE):
        w, c, st = solve(torch, handle, h, s, prec)
        g = workloads.gates(h, s, w, c, wref)
        print(prec, st, handle.refine_check(), g, flush=True)
        assert st == [0, 0, 0, 0], (prec, st)
        assert workloads.passes(g), (prec, g)
        if prec == xtbeig.FP32_REFINE:
            q = handle.refine_check()
            assert q[0] <= 1.0 and q[1] <= 1.0, q


@pytest.mark.parametrize("n", [1204])
def test_independent_h_ill_conditioned(handle, torch_cuda, n):
    """H independent of S (the configuration's literal recipe) with cond_2(S) = 1e4: the generalized eigenvalues grow
    to ||H|| / lambda_min(S); FP64 on the device and the host-array mixed call (refined, or redone in FP64 when the
    check fails) both pass the scaled gates."""
    torch = torch_cuda
    handle.set_tuning(0, 2)
    h, s, _ = workloads.generalized_pair_cond(n, workloads.SEED_BASE + 1, 1e4, physical=False)
    wref = sl.eigh(h, s, eigvals_only=True)
    print(f"cond2(S)={cond2(s):.4e} max|w|={np.max(np.abs(wref)):.3e} ||H||={np.linalg.norm(h, 2):.3f}", flush=True)
    w, c, st = solve(torch, handle, h, s, xtbeig.FP64)
    assert st == [0, 0, 0, 0]
    g = workloads.gates(h, s, w, c, wref, scaled=True)
    assert workloads.passes(g), g
    w2, c2, u, info = xtbeig.dsygvd(h, s, mixed=True)
    assert info == 0
    g2 = workloads.gates(h, s, w2, c2, wref, scaled=True)
    assert workloads.passes(g2), g2


@pytest.mark.parametrize("n", [588, 1204])
def test_near_degenerate_pairs_mixed(handle, torch_cuda, n):
    """The density configuration's spectrum (8 near-degenerate pairs, gaps 1e-10 .. 1e-4, plus the repeated and
    1e-9-split pairs) as a generalized spectrum: the FP32 solve mixes such pairs; the cluster step of the refinement
    resolves them."""
    torch = torch_cuda
    handle.set_tuning(0, 2)
    h, s, lam = workloads.density_problem(n, workloads.SEED_BASE + 4)
    wref = sl.eigh(h, s, eigvals_only=True)
    w, c, st = solve(torch, handle, h, s, xtbeig.FP32_REFINE)
    g = workloads.gates(h, s, w, c, wref)
    print(st, handle.refine_check(), g, flush=True)
    assert st == [0, 0, 0, 0], st
    assert workloads.passes(g), g
    assert np.max(np.abs(w - lam)) < 1e-12 * np.linalg.norm(h, 2)


def test_check_flags_unconverged_refinement(handle, torch_cuda):
    """One refinement step from FP32 at cond_2(S) = 1e4 does not reach the gates: the check says so (status word 3,
    ratios above 1), matching the host gates."""
    torch = torch_cuda
    n = 600
    h, s, _ = workloads.generalized_pair_cond(n, workloads.SEED_BASE + 1, 1e4, physical=True)
    wref = sl.eigh(h, s, eigvals_only=True)
    handle.set_tuning(0, 1)
    try:
        w, c, st = solve(torch, handle, h, s, xtbeig.FP32_REFINE)
        q = handle.refine_check()
    finally:
        handle.set_tuning(0, 2)
    g = workloads.gates(h, s, w, c, wref)
    assert not workloads.passes(g), g
    assert st[:3] == [0, 0, 0] and st[3] != 0, st
    assert max(q) > 1.0, q


def test_scc_restart_on_failed_check(torch_cuda):
    """An FP32 SCC cycle whose refined pairs fail the check is redone in FP64: fp64_restart = 1, the cycle reports
    FP64, and its result is bit-identical to a plain FP64 solve."""
    torch = torch_cuda
    n = 600
    h, s, _ = workloads.generalized_pair_cond(n, workloads.SEED_BASE + 1, 1e4, physical=True)
    hd, sd = dev(torch, h), dev(torch, s)
    hnd = xtbeig.Handle(0)
    try:
        hnd.set_tuning(0, 1)
        hnd.set_scc_precision(1e-3, 50, 1)
        hnd.scc_begin()
        w = torch.zeros(n, dtype=torch.float64, device="cuda")
        c = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        used = hnd.scc_sygvd(hd, sd, w, c, 1.0)
        settings, counters, rc = hnd.get_scc_precision()
        assert used == xtbeig.FP64
        assert counters["fp64_restart"] == 1 and counters["fp32_cycles"] == 0 and counters["scc_cycles"] == 1
        w64 = torch.zeros_like(w)
        c64 = torch.zeros_like(c)
        hnd.sygvd(hd, sd, w64, c64, xtbeig.FP64)
        assert hnd.synchronize() == [0, 0, 0, 0]
        assert torch.equal(w, w64) and torch.equal(c, c64)
        # the sequence stays FP64 after the restart
        used = hnd.scc_sygvd(hd, sd, w, c, 1.0)
        assert used == xtbeig.FP64
    finally:
        hnd.close()


def test_graph_follows_refinement_steps(torch_cuda):
    """The graph cache keys on the refinement step count: replays after xtbeig_set_tuning(0, 0) and (0, 3) are
    bit-identical to eager solves with the same settings."""
    torch = torch_cuda
    n = 300
    h, s = workloads.generalized_pair(n, workloads.SEED_BASE + 1)
    hd, sd = dev(torch, h), dev(torch, s)
    hnd = xtbeig.Handle(0)
    try:
        for steps in (2, 0, 3, 2):
            hnd.set_tuning(0, steps)
            outs = []
            for graph in (False, True, True):
                w = torch.zeros(n, dtype=torch.float64, device="cuda")
                c = torch.zeros((n, n), dtype=torch.float64, device="cuda")
                if graph:
                    hnd.sygvd_graph(hd, sd, w, c, xtbeig.FP32_REFINE)
                else:
                    hnd.sygvd(hd, sd, w, c, xtbeig.FP32_REFINE)
                st = hnd.synchronize()
                assert st[:3] == [0, 0, 0], st
                outs.append((w, c))
            for w, c in outs[1:]:
                assert torch.equal(w, outs[0][0]) and torch.equal(c, outs[0][1]), steps
    finally:
        hnd.close()
