"""Failure paths and the remaining entry points of the sytrd.c symbol set.

  * A non-converged secular equation is reported as LAPACK info, never aborted on (the reference propagates a failed
    dlaed4 as info, offset by the merge's first row: Stages/ParallelDivideAndConquer.swift:70-71): the test setting
    xtbeig_set_test_limits caps the secular iterations at 0, so that the roots give up after their starting guess
    (every kernel still runs to its end; the values stay finite and ordered) and the status word and info say so.
  * xtb_dsymmetrize_pad (Exports.swift:117-138) against numpy's mirror of the upper triangle, padded leading
    dimension.
  * xtb_sstedc (Exports.swift:103-115) against the FP64 tridiagonal eigensolver of scipy.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sl

import workloads
import xtbeig

pytestmark = pytest.mark.gpu


def tridiag(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, n), rng.uniform(0.1, 1, n - 1)


def test_device_stedc_reports_unconverged_roots(handle, torch_cuda):
    torch = torch_cuda
    n = 300
    d0, e0 = tridiag(n, 7)
    z = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    handle.set_test_limits(0)
    try:
        d = torch.from_numpy(d0.copy()).cuda()
        e = torch.from_numpy(np.append(e0, 0.0)).cuda()
        handle.dstedc(d, e, z)
        st = handle.synchronize()
    finally:
        handle.set_test_limits(-1)
    assert st[0] == 0 and st[1] == 0 and st[3] == 0
    assert 0 < st[2] <= n, st
    assert np.all(np.isfinite(d.cpu().numpy()))
    # the same call with the solver's own limit converges again
    d = torch.from_numpy(d0.copy()).cuda()
    e = torch.from_numpy(np.append(e0, 0.0)).cuda()
    handle.dstedc(d, e, z)
    assert handle.synchronize() == [0, 0, 0, 0]
    wref = sl.eigh_tridiagonal(d0, e0, eigvals_only=True)
    assert np.max(np.abs(d.cpu().numpy() - wref)) < 1e-13


def test_host_calls_return_info_on_unconverged_roots(torch_cuda):
    L = xtbeig.lib()
    n = 200
    d0, e0 = tridiag(n, 8)
    h, s = workloads.generalized_pair(n, workloads.SEED_BASE + 1)
    assert L.xtbeig_set_test_limits(None, 0) == 0
    try:
        _, _, info = xtbeig.stedc(d0, e0)
        assert 0 < info <= n, info
        _, _, info = xtbeig.syevd(h)
        assert 0 < info <= n, info
        _, _, _, info = xtbeig.dsygvd(h, s)
        assert 0 < info <= n, info
    finally:
        assert L.xtbeig_set_test_limits(None, -1) == 0
    w, c, u, info = xtbeig.dsygvd(h, s)
    assert info == 0
    assert workloads.passes(workloads.gates(h, s, w, c, sl.eigh(h, s, eigvals_only=True)))


def test_scc_cycle_reports_a_failed_fp64_redo(torch_cuda):
    """An FP32 cycle whose divide and conquer fails is redone in FP64 (fp64_restart = 1); when that fails as well the
    call returns an error instead of results."""
    torch = torch_cuda
    n = 200
    h, s = workloads.generalized_pair(n, workloads.SEED_BASE + 1)
    hd = torch.from_numpy(np.ascontiguousarray(h.T)).cuda()
    sd = torch.from_numpy(np.ascontiguousarray(s.T)).cuda()
    w = torch.zeros(n, dtype=torch.float64, device="cuda")
    c = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    hnd = xtbeig.Handle(0)
    try:
        hnd.set_scc_precision(1e-3, 50, 1)
        hnd.scc_begin()
        hnd.set_test_limits(0)
        with pytest.raises(RuntimeError, match="FP64 eigensolve failed"):
            hnd.scc_sygvd(hd, sd, w, c, 1.0)
        _, counters, _ = hnd.get_scc_precision()
        assert counters["fp64_restart"] == 1
        hnd.set_test_limits(-1)
        hnd.scc_begin()
        assert hnd.scc_sygvd(hd, sd, w, c, 1.0) == xtbeig.FP32_REFINE
    finally:
        hnd.close()


@pytest.mark.parametrize("n,pad", [(1, 0), (37, 3), (600, 64), (1204, 8)])
def test_symmetrize_pad(torch_cuda, n, pad):
    rng = np.random.default_rng(n)
    src = np.asfortranarray(rng.standard_normal((n + 5, n)))  # lds = n + 5; the lower triangle is garbage
    dst = xtbeig.dsymmetrize_pad(src, ldd=n + pad, n=n)
    up = np.triu(src[:n, :n])
    ref = up + np.triu(up, 1).T
    assert np.array_equal(dst, ref)


@pytest.mark.parametrize("n", [1, 2, 17, 300, 1204])
def test_sstedc_against_fp64(torch_cuda, n):
    d0, e0 = tridiag(n, 3 + n)
    d32, e32 = d0.astype(np.float32), e0.astype(np.float32)
    w, z, info = xtbeig.stedc(d32, e32, dtype=np.float32)
    assert info == 0 and w.dtype == np.float32
    if n == 1:
        assert w[0] == d32[0] and z[0, 0] == 1
        return
    wref = sl.eigh_tridiagonal(d32.astype(np.float64), e32.astype(np.float64), eigvals_only=True)
    scale = max(np.max(np.abs(wref)), 1.0)
    assert np.max(np.abs(w.astype(np.float64) - wref)) < 1e-6 * scale
    t = np.diag(d32.astype(np.float64)) + np.diag(e32.astype(np.float64), 1) + np.diag(e32.astype(np.float64), -1)
    z64 = z.astype(np.float64)
    assert np.max(np.abs(t @ z64 - z64 * w.astype(np.float64))) < 1e-5 * scale
    assert np.max(np.abs(z64.T @ z64 - np.eye(n))) < 1e-5
