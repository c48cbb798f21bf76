"""The generalized solve at the workload sizes (n = 1204; n = 1788 with the band-32 and band-64 configurations and
with the built-in band; band 48 at n = 1204) and at sizes on either side of the 64-row tiles and of n = 1204, in both precisions, against host
LAPACK (scipy eigh) with the FP64 gates: eig err <= 1e-12 ||H||, residual <= 1e-13, ||C^T S C - I|| <= 1e-12 n."""
import numpy as np
import pytest
import scipy.linalg as sl

import workloads
import xtbeig

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(np.asfortranarray(a).T)).cuda()


def host(t):
    return np.asfortranarray(t.cpu().numpy().T)


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("n,band", [(1204, 0), (1788, 32), (1788, 0), (63, 0), (65, 0), (127, 0), (129, 0),
                                    (257, 0), (1203, 0), (1205, 0), (600, 16), (700, 32),
                                    (1788, 64), (1204, 48), (600, 64), (100, 64)])
def test_sygvd_sizes(handle, torch_cuda, n, band, mixed):
    torch = torch_cuda
    h, s = workloads.generalized_pair(n, 1001 + n)
    print(f"n={n} band={band} mixed={mixed}", flush=True)
    handle.set_tuning(band, 2)
    try:
        H, S = dev(torch, h), dev(torch, s)
        W = torch.zeros(n, dtype=torch.float64, device="cuda")
        C = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        handle.sygvd(H, S, W, C, xtbeig.FP32_REFINE if mixed else xtbeig.FP64)
        assert handle.synchronize() == [0, 0, 0, 0]
    finally:
        handle.set_tuning(0, 2)
    g = workloads.gates(h, s, W.cpu().numpy(), host(C), sl.eigh(h, s, eigvals_only=True))
    assert workloads.passes(g), (n, band, mixed, g)


def test_sygvd_refuses_sizes_past_the_limit(handle, torch_cuda):
    # n - band > 2048 is refused on the host before any work is enqueued; the handle keeps working
    torch = torch_cuda
    n = 2100
    H = torch.eye(n, dtype=torch.float64, device="cuda")
    S = H.clone()
    W = torch.zeros(n, dtype=torch.float64, device="cuda")
    C = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="exceeds the largest supported size"):
        handle.sygvd(H, S, W, C, xtbeig.FP64)
    assert handle.synchronize() == [0, 0, 0, 0]
    m = 131
    h, s = workloads.generalized_pair(m, 5)
    H2, S2 = dev(torch, h), dev(torch, s)
    W2 = torch.zeros(m, dtype=torch.float64, device="cuda")
    C2 = torch.zeros((m, m), dtype=torch.float64, device="cuda")
    handle.sygvd(H2, S2, W2, C2, xtbeig.FP32_REFINE)
    assert handle.synchronize() == [0, 0, 0, 0]
    g = workloads.gates(h, s, W2.cpu().numpy(), host(C2), sl.eigh(h, s, eigvals_only=True))
    assert workloads.passes(g), g


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("uplo", ["U", "L"])
def test_host_sygvd_staged_and_strided(torch_cuda, uplo, mixed):
    """The LAPACK-style host-array call on matrices large enough (2.9 MB) and strided enough (lda = n + 7) to go
    through the handle's pinned staging buffer: results pass the gates, the padding rows past n come back untouched,
    and b holds the Cholesky factor in the referenced triangle with the other one untouched."""
    n, ld = 600, 607
    h, s = workloads.generalized_pair(n, 77)
    a = np.full((ld, n), 7.0, order="F")
    b = np.full((ld, n), 9.0, order="F")
    a[:n] = h
    b[:n] = s
    w = np.zeros(n)
    lw, liw = 1 + 6 * n + 2 * n * n, 3 + 5 * n
    work = np.zeros(lw)
    iwork = np.zeros(liw, dtype=np.int32)
    info = xtbeig.ctypes.c_int(0)
    f = xtbeig.lib().xtbeig_dsygvd_mixed if mixed else xtbeig.lib().xtbeig_dsygvd
    f(xtbeig._i(1), b"V", uplo.encode(), xtbeig._i(n), xtbeig._ptr(a), xtbeig._i(ld), xtbeig._ptr(b),
      xtbeig._i(ld), xtbeig._ptr(w), xtbeig._ptr(work), xtbeig._i(lw), xtbeig._ptr(iwork), xtbeig._i(liw),
      xtbeig.ctypes.byref(info))
    assert info.value == 0, xtbeig.last_error()
    assert np.all(a[n:] == 7.0) and np.all(b[n:] == 9.0)
    g = workloads.gates(h, s, w, np.asfortranarray(a[:n]), sl.eigh(h, s, eigvals_only=True))
    assert workloads.passes(g), g
    f_ = b[:n]
    if uplo == "U":
        u = np.triu(f_)
        assert np.max(np.abs(u.T @ u - s)) < 1e-14 * n * np.linalg.norm(s, 2)
        assert np.array_equal(np.tril(f_, -1), np.tril(s, -1))
    else:
        lo = np.tril(f_)
        assert np.max(np.abs(lo @ lo.T - s)) < 1e-14 * n * np.linalg.norm(s, 2)
        assert np.array_equal(np.triu(f_, 1), np.triu(s, 1))
