"""Adversarial inputs through the device generalized solve (S = I, so the pins are the standard eigenproblem's):
degenerate groups of 3 and 8, a 1e-8 cluster, two values, a diagonal, the identity, zero, a tridiagonal, matrices
scaled by 1e150 and 1e-150, a graded spectrum and a block-diagonal matrix with equal blocks.

Pins: host LAPACK (scipy.linalg.eigh) with the contract's gates. FP64 must pass every case. FP32 + refinement must
either pass or report it cannot (status word 3, the refinement check): spectra that FP32 cannot resolve (a cluster
narrower than FP32's rounding, eigenvalues below it, clusters wider than the Rayleigh-Ritz step's 32) are redone in
FP64 by the host and SCC entry points on that status, and no case may return failing pairs with a clean status.
Matrices far outside FP32's range are scaled by a power of two for the FP32 solve (1e150 once gave non-finite values
there)."""
import numpy as np
import pytest
import scipy.linalg as sl

import xtbeig

pytestmark = pytest.mark.gpu

N = 300


def cases():
    rng = np.random.default_rng(5)
    n = N

    def with_spec(lam):
        q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        a = (q * lam) @ q.T
        return (a + a.T) / 2

    out = {}
    out["mult3"] = with_spec(np.sort(np.tile(rng.uniform(-1, 1, n // 3 + 1), 3)[:n]))
    out["mult8"] = with_spec(np.sort(np.tile(rng.uniform(-1, 1, n // 8 + 1), 8)[:n]))
    out["cluster1e-8"] = with_spec(1.0 + 1e-8 * rng.uniform(0, 1, n))
    out["two_values"] = with_spec(np.sort(np.where(rng.uniform(size=n) < 0.5, -1.0, 1.0)))
    out["diag"] = np.diag(rng.uniform(-1, 1, n))
    out["identity"] = np.eye(n)
    out["zero"] = np.zeros((n, n))
    e = rng.uniform(-1, 1, n - 1)
    out["tridiag"] = np.diag(rng.uniform(-1, 1, n)) + np.diag(e, 1) + np.diag(e, -1)
    out["scaled1e150"] = with_spec(rng.uniform(-1, 1, n)) * 1e150
    out["scaled1e-150"] = with_spec(rng.uniform(-1, 1, n)) * 1e-150
    out["graded"] = with_spec(np.sort(np.logspace(-12, 0, n)))
    blk = with_spec(rng.uniform(-1, 1, n))[: n // 2, : n // 2]
    bd = np.zeros((n, n))
    bd[: n // 2, : n // 2] = blk
    bd[n // 2:, n // 2:] = blk
    out["blockdiag_dup"] = bd
    return out


CASES = cases()


@pytest.mark.parametrize("name", list(CASES))
def test_adversarial(handle, torch_cuda, name):
    torch = torch_cuda
    a = CASES[name]
    n = a.shape[0]
    wref = sl.eigh(a, eigvals_only=True)
    nh = max(np.linalg.norm(a, 2), np.finfo(float).tiny)
    for prec in (xtbeig.FP64, xtbeig.FP32_REFINE):
        w = torch.zeros(n, dtype=torch.float64, device="cuda")
        c = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        hd = torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
        sd = torch.eye(n, dtype=torch.float64, device="cuda")
        handle.sygvd(hd, sd, w, c, prec)
        st = handle.synchronize()
        ww, cc = w.cpu().numpy(), c.cpu().numpy().T
        eig = np.max(np.abs(ww - wref)) / nh
        res = np.linalg.norm(a @ cc - cc * ww) / (n * nh)
        orth = np.max(np.abs(cc.T @ cc - np.eye(n)))
        ok = eig <= 1e-12 and res <= 1e-13 and orth <= 1e-12 * n
        print(name, prec, st, f"eig {eig:.1e} res {res:.1e} orth {orth:.1e}", flush=True)
        assert st[:3] == [0, 0, 0], (name, prec, st)
        if prec == xtbeig.FP64:
            assert st[3] == 0 and ok, (name, eig, res, orth)
        else:
            assert ok or st[3] != 0, (name, eig, res, orth, st)


def fp32_near_underflow_cases():
    rng = np.random.default_rng(21)
    n = N
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    out = {"rank1": np.ones((n, n))}  # a 299-fold zero eigenvalue: the chase's bulges decay below FP32's squares
    h = n // 2
    a = np.zeros((n, n))
    qa, _ = np.linalg.qr(rng.standard_normal((h, h)))
    qb, _ = np.linalg.qr(rng.standard_normal((n - h, n - h)))
    a[:h, :h] = (qa * rng.uniform(-1, 1, h)) @ qa.T
    a[h:, h:] = 1e-21 * (qb * rng.uniform(-1, 1, n - h)) @ qb.T  # panels whose columns' squares underflow in FP32
    out["block_1e-21"] = (a + a.T) / 2
    return out


FP32_CASES = fp32_near_underflow_cases()


@pytest.mark.parametrize("band", [0, 12, 32, 48])
@pytest.mark.parametrize("name", list(FP32_CASES))
def test_fp32_reflectors_stay_orthogonal(handle, torch_cuda, name, band):
    """The FP32 standard solve where Householder vectors get so small that their squares underflow in FP32 (the
    bulges of a massively degenerate spectrum, a block 1e-21 below the rest): the reflectors are formed in FP64
    there, so the eigenvectors stay orthogonal to FP32 rounding (a sum of squares that lost terms to underflow once
    gave an orthogonality error of 0.24). Eigenvalues to FP32 rounding of ||A||. Band widths: the built-in 8 (fused
    panels, lane chase), 12 (general-path panels, register bulges), 32 (lane chase with one wavefront per bulge),
    48 (block bulges)."""
    torch = torch_cuda
    a = FP32_CASES[name]
    n = a.shape[0]
    A = torch.from_numpy(np.ascontiguousarray(a.T)).cuda().float()
    w = torch.zeros(n, dtype=torch.float64, device="cuda")  # the eigenvalues come back in FP64
    z = torch.zeros((n, n), dtype=torch.float32, device="cuda")
    handle.set_tuning(band, 2)
    try:
        handle.ssyevd(A, w, z)
        assert handle.synchronize() == [0, 0, 0, 0]
    finally:
        handle.set_tuning(0, 2)
    ww, zz = w.double().cpu().numpy(), z.double().cpu().numpy().T
    ref = np.linalg.eigvalsh(a)
    nrm = np.abs(ref).max()
    print(name, band, np.max(np.abs(ww - ref)) / nrm, np.max(np.abs(zz.T @ zz - np.eye(n))), flush=True)
    assert np.max(np.abs(ww - ref)) < 1e-5 * nrm
    assert np.max(np.abs(zz.T @ zz - np.eye(n))) < 1e-5 * n ** 0.5
