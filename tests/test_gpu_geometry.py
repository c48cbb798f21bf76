"""The steps either side of the reduction in xTB's per-cycle eigensolve, and the solve on a cached geometry.

Pins:
  * xtb_dcongruence_upper against the host LAPACK's dsygst(1, 'U') (which forms U^-T H U^-1 = X^T H X) on the upper
    triangle; the strictly lower triangle of Z is left bit for bit as it was (congruence.c:1-8 writes the upper
    triangle only);
  * xtb_dtrmm_back against the host LAPACK's dtrsm('L', 'U', 'N', 'N') (U^-1 C = X C) on the first m columns; the
    other columns untouched;
  * the device twins on a geometry set once (xtbeig_dev_set_geometry) against the same host results;
  * xtbeig_dev_sygvd_geometry over three cycles of one geometry: bit-identical to xtbeig_dev_sygvd's eigenvalues and
    first m eigenvectors (the same kernels on the same X), and the reference's gates on the m columns
    (eigbench.f90:189-204: max |H C - S C Lambda| and max |C^T S C - I|);
  * xtbeig_dev_dsyevd_partial: bit-identical to the first m columns of xtbeig_dev_dsyevd.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import host_lapack as hl
import workloads
import xtbeig

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda()


def geometry(n, seed):
    h, s = workloads.generalized_pair(n, seed)
    u, info = hl.dpotrf_upper(s)
    assert info == 0
    u = np.triu(u)
    x = np.triu(sl.solve_triangular(u, np.eye(n)))
    return h, s, u, x


@pytest.mark.parametrize("n", [64, 600, 1204])
def test_host_congruence_and_back_solve(torch_cuda, n):
    h, s, u, x = geometry(n, workloads.SEED_BASE + 1)
    ref, info = hl.dsygst_upper(h, u)
    assert info == 0
    rng = np.random.default_rng(n)
    z0 = rng.standard_normal((n, n))
    z = xtbeig.dcongruence_upper(h, x, z0)
    iu = np.triu_indices(n)
    il = np.tril_indices(n, -1)
    scale = np.max(np.abs(ref[iu]))
    assert np.max(np.abs(z[iu] - ref[iu])) < 1e-12 * scale
    assert np.array_equal(z[il], z0[il])
    m = n // 2
    c0 = rng.standard_normal((n, n))
    c = xtbeig.dtrmm_back(x, c0, m)
    cref = hl.dtrsm_left_upper(u, c0[:, :m])
    assert np.max(np.abs(c[:, :m] - cref)) < 1e-12 * np.max(np.abs(cref))
    assert np.array_equal(c[:, m:], c0[:, m:])


def test_device_congruence_and_back_solve_on_geometry(handle, torch_cuda):
    torch = torch_cuda
    n = 600
    h, s, u, x = geometry(n, workloads.SEED_BASE + 1)
    sd, hd = dev(torch, s), dev(torch, h)
    handle.set_geometry(sd)
    assert handle.synchronize() == [0, 0, 0, 0]
    z0 = np.random.default_rng(3).standard_normal((n, n))
    zd = dev(torch, z0)
    handle.congruence(hd, zd)
    m = 100
    c0 = np.random.default_rng(4).standard_normal((n, n))
    cd = dev(torch, c0)
    handle.back_solve(cd, m)
    assert handle.synchronize() == [0, 0, 0, 0]
    z = zd.cpu().numpy().T
    ref, _ = hl.dsygst_upper(h, u)
    iu = np.triu_indices(n)
    il = np.tril_indices(n, -1)
    assert np.max(np.abs(z[iu] - ref[iu])) < 1e-12 * np.max(np.abs(ref[iu]))
    assert np.array_equal(z[il], z0[il])
    c = cd.cpu().numpy().T
    cref = hl.dtrsm_left_upper(u, c0[:, :m])
    assert np.max(np.abs(c[:, :m] - cref)) < 1e-12 * np.max(np.abs(cref))
    assert np.array_equal(c[:, m:], c0[:, m:])


def eigbench_gates(h, s, w, c):
    """The reference's report quantities on the columns of c (eigbench.f90:189-204)."""
    r = h @ c - (s @ c) * w[: c.shape[1]]
    o = c.T @ s @ c - np.eye(c.shape[1])
    return float(np.max(np.abs(r))), float(np.max(np.abs(o)))


@pytest.mark.parametrize("prec", [xtbeig.FP64, xtbeig.FP32_REFINE])
def test_sygvd_on_cached_geometry(torch_cuda, prec):
    torch = torch_cuda
    n = 1204
    h0, s = workloads.generalized_pair(n, workloads.SEED_BASE + 2)
    sd = dev(torch, s)
    hnd = xtbeig.Handle(0)
    try:
        hnd.set_geometry(sd)
        assert hnd.synchronize() == [0, 0, 0, 0]
        rng = np.random.default_rng(7)
        for cycle, m in enumerate((n, n // 2, 300)):
            d = rng.standard_normal((n, n))
            hk = h0 + 1e-3 * (0.5 * (d + d.T)) / n
            hd = dev(torch, hk)
            w = torch.zeros(n, dtype=torch.float64, device="cuda")
            c = torch.zeros((m, n), dtype=torch.float64, device="cuda")  # column-major n x m, ld n
            hnd.sygvd_geometry(hd, w, c, m, prec)
            wf = torch.zeros(n, dtype=torch.float64, device="cuda")
            cf = torch.zeros((n, n), dtype=torch.float64, device="cuda")
            hnd.sygvd(hd, sd, wf, cf, prec)
            st = hnd.synchronize()
            assert st == [0, 0, 0, 0], (cycle, st)
            assert torch.equal(w, wf), cycle
            assert torch.equal(c, cf[:m]), cycle
            wn, cn = w.cpu().numpy(), c.cpu().numpy().T
            wref = sl.eigh(hk, s, eigvals_only=True)
            nh = np.linalg.norm(hk, 2)
            assert np.max(np.abs(wn - wref)) <= 1e-12 * nh
            res, orth = eigbench_gates(hk, s, wn, cn)
            assert res <= 1e-12 * nh and orth <= 1e-12 * n, (cycle, res, orth)
    finally:
        hnd.close()


@pytest.mark.parametrize("m", [1, 64, 602])
def test_dsyevd_partial(handle, torch_cuda, m):
    torch = torch_cuda
    n = 1204
    a, lam = workloads.standard_matrix(n, 11)
    ad = dev(torch, a)
    w = torch.zeros(n, dtype=torch.float64, device="cuda")
    z = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    handle.dsyevd_partial(ad, w, z, m)
    wf = torch.zeros_like(w)
    zf = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    handle.dsyevd(ad, wf, zf)
    assert handle.synchronize() == [0, 0, 0, 0]
    assert torch.equal(w, wf) and torch.equal(z, zf[:m])
    assert np.max(np.abs(w.cpu().numpy() - lam)) < 1e-12 * np.max(np.abs(lam))
