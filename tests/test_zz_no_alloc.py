"""The allocation-free Cholesky + inverse (xtbeig_dev_potrf_inverse_noalloc) against xtbeig_dev_potrf_inverse,
through the C ABI: bit-identical U and X on two calls in a row (the second smaller), no change in the device's free
memory across the allocation-free calls, and the size limit refused with an error."""
import ctypes

import numpy as np
import pytest

import workloads
import xtbeig

pytestmark = pytest.mark.gpu


def overlap(n, seed):
    return workloads.overlap_matrix(n, np.random.default_rng(seed))


def run(torch, L, fn, hnd, s):
    n = s.shape[0]
    sd = torch.from_numpy(np.ascontiguousarray(s.T)).cuda()
    u = torch.full((n, n), float("nan"), dtype=torch.float64, device="cuda")
    x = torch.full((n, n), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = getattr(L, fn)(hnd._h, n, ctypes.c_void_p(sd.data_ptr()), n, ctypes.c_void_p(u.data_ptr()), n,
                        ctypes.c_void_p(x.data_ptr()), n)
    assert rc == 0, xtbeig.last_error()
    st = hnd.synchronize()
    return u, x, st


def test_noalloc_matches_potrf_inverse(torch_cuda):
    torch = torch_cuda
    L = xtbeig.lib()
    hnd = xtbeig.Handle(0)
    try:
        # every kernel of the call loaded first (the runtime loads code objects on first use, which takes device
        # memory), outside the measured region: n = 200 runs the diagonal blocks, panels, trailing updates and inverse
        run(torch, L, "xtbeig_dev_potrf_inverse_noalloc", hnd, overlap(200, 1))
        mats = [overlap(1204, 11), overlap(600, 12)]
        ins = [torch.from_numpy(np.ascontiguousarray(s.T)).cuda() for s in mats]
        outs = [torch.zeros_like(t) for t in ins for _ in range(2)]
        work = [t.clone() for t in ins]  # destroyed by the factorization
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        got = []
        for q, s in enumerate(mats):
            n = s.shape[0]
            sd = work[q]
            u, x = outs[2 * q], outs[2 * q + 1]
            torch.cuda.synchronize()
            free_a = torch.cuda.mem_get_info()[0]
            rc = L.xtbeig_dev_potrf_inverse_noalloc(hnd._h, n, ctypes.c_void_p(sd.data_ptr()), n,
                                                    ctypes.c_void_p(u.data_ptr()), n, ctypes.c_void_p(x.data_ptr()),
                                                    n)
            assert rc == 0, xtbeig.last_error()
            st = hnd.synchronize()
            free_b = torch.cuda.mem_get_info()[0]
            assert st == [0, 0, 0, 0], st
            assert free_b == free_a, (free_a, free_b)
        assert torch.cuda.mem_get_info()[0] == free0
        got = [(outs[2 * q], outs[2 * q + 1]) for q in range(len(mats))]
        for q, s in enumerate(mats):
            u_ref, x_ref, st = run(torch, L, "xtbeig_dev_potrf_inverse", hnd, s)
            assert st == [0, 0, 0, 0], st
            u, x = got[q]
            assert torch.equal(u, u_ref) and torch.equal(x, x_ref), s.shape
            un, xn = u.cpu().numpy().T, x.cpu().numpy().T
            n = s.shape[0]
            assert np.max(np.abs(un.T @ un - s)) < 1e-13 * n
            assert np.max(np.abs(xn @ un - np.eye(n))) < 1e-11
    finally:
        hnd.close()


def test_noalloc_rejects_oversize(torch_cuda):
    torch = torch_cuda
    L = xtbeig.lib()
    hnd = xtbeig.Handle(0)
    try:
        n = 2049
        buf = torch.zeros(8, dtype=torch.float64, device="cuda")
        p = ctypes.c_void_p(buf.data_ptr())
        rc = L.xtbeig_dev_potrf_inverse_noalloc(hnd._h, n, p, n, p, n, p, n)
        assert rc != 0
        assert "2048" in xtbeig.last_error()
        assert hnd.synchronize() == [0, 0, 0, 0]
    finally:
        hnd.close()
