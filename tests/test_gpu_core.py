"""Stage-by-stage parity of the HIP kernels against the host-LAPACK oracle (scipy's OpenBLAS) and numpy."""
import numpy as np
import pytest
import scipy.linalg as sl

import sbr
import workloads
import xtbeig

pytestmark = pytest.mark.gpu


def dev(torch, a, dtype=None):
    a = np.asfortranarray(a)
    t = torch.from_numpy(np.ascontiguousarray(a.T if a.ndim == 2 else a)).cuda()
    if dtype is not None:
        t = t.to(dtype)
    return t  # row-major storage of a^T == column-major a


def host(t):
    a = t.cpu().numpy()
    return np.asfortranarray(a.T) if a.ndim == 2 else a


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("prec", ["d", "s"])
def test_gemm(handle, torch_cuda, ta, tb, prec):
    torch = torch_cuda
    rng = np.random.default_rng(7)
    m, n, k = 131, 77, 203
    a = rng.standard_normal((m, k) if ta == "N" else (k, m))
    b = rng.standard_normal((k, n) if tb == "N" else (n, k))
    c = rng.standard_normal((m, n))
    dt = torch.float64 if prec == "d" else torch.float32
    A, B, C = dev(torch, a, dt), dev(torch, b, dt), dev(torch, c, dt)
    handle.gemm(ta, tb, m, n, k, 1.5, A, a.shape[0], B, b.shape[0], -0.5, C, m)
    handle.synchronize()
    opa = a if ta == "N" else a.T
    opb = b if tb == "N" else b.T
    ref = 1.5 * opa @ opb - 0.5 * c
    tol = 1e-13 if prec == "d" else 2e-5
    assert np.max(np.abs(host(C) - ref)) <= tol * np.max(np.abs(ref))


def test_potrf_inverse(handle, torch_cuda):
    torch = torch_cuda
    n = 300
    s = workloads.overlap_matrix(n, np.random.default_rng(3))
    S = dev(torch, s)
    U = torch.zeros_like(S)
    X = torch.zeros_like(S)
    handle.potrf_inverse(S, U, X)
    assert handle.synchronize() == [0, 0, 0, 0]
    u, x = host(U), host(X)
    ref = sl.cholesky(s, lower=False)
    assert np.max(np.abs(u - ref)) < 1e-13 * np.max(np.abs(ref))
    assert np.max(np.abs(np.tril(u, -1))) == 0 and np.max(np.abs(np.tril(x, -1))) == 0
    assert np.max(np.abs(x @ u - np.eye(n))) < 1e-12


@pytest.mark.parametrize("bad", [0, 70, 263])
def test_potrf_not_positive_definite(handle, torch_cuda, bad):
    # LAPACK's dpotrf convention: info = order of the first leading minor that is not positive definite
    torch = torch_cuda
    n = 300
    s = workloads.overlap_matrix(n, np.random.default_rng(3))
    s[bad, bad] = -1.0
    S = dev(torch, s)
    U = torch.zeros_like(S)
    X = torch.zeros_like(S)
    handle.potrf_inverse(S, U, X)
    assert handle.synchronize()[0] == bad + 1  # reading the status also clears it


@pytest.mark.parametrize("prec,b,n", [("d", 32, 260), ("d", 8, 131), ("s", 16, 200), ("d", 32, 33), ("d", 4, 9),
                                      ("d", 16, 700), ("s", 8, 1300), ("d", 8, 1100), ("d", 16, 1700),
                                      # the widest panels of the general path
                                      ("d", 64, 300), ("s", 48, 400), ("d", 48, 1100), ("s", 64, 1788)])
def test_reduce_to_band(handle, torch_cuda, prec, b, n):
    torch = torch_cuda
    a, lam = workloads.standard_matrix(n, 11)
    dt = torch.float64 if prec == "d" else torch.float32
    A = dev(torch, a, dt)
    V1 = torch.zeros((n, n), dtype=dt, device="cuda")
    T1 = torch.zeros((n, b), dtype=dt, device="cuda")
    AB = torch.zeros(xtbeig.band_elems(n, b), dtype=dt, device="cuda")
    handle.reduce_to_band(A, b, V1, T1, AB)
    handle.synchronize()
    band = xtbeig.band_to_dense(AB.double().cpu().numpy(), n, b)
    assert np.all(np.abs(np.tril(band, -b - 1)) == 0)
    wb = np.linalg.eigvalsh(band)
    tol = 1e-13 if prec == "d" else 3e-6
    assert np.max(np.abs(wb - lam)) < tol
    # same algorithm as the oracle's restatement of BandForm.swift: same band to rounding (in FP32 up to the sign of
    # a reflector whose leading element is within rounding of zero, which flips a row and column of the band: the
    # magnitudes are compared, and Q1^T A Q1 = band below pins the signs)
    oband, _, _ = sbr.reduce_to_band(a, b)
    if prec == "d":
        assert np.max(np.abs(band - oband)) < 1e-12 * n
    else:
        dm = np.abs(np.abs(band) - np.abs(oband))
        assert np.max(dm) < 2e-5 * n, (np.unravel_index(np.argmax(dm), dm.shape), np.max(dm),
                                       np.max(np.abs(band - oband)))
    # back-transformation through the panels: Q1 = (back_band of I), Q1^T A Q1 = band
    Z = torch.eye(n, dtype=dt, device="cuda")
    handle.back_band(n, b, V1, T1, Z, n)
    handle.synchronize()
    q1 = host(Z).astype(np.float64)
    assert np.max(np.abs(q1.T @ q1 - np.eye(n))) < (1e-13 if prec == "d" else 3e-6) * n
    assert np.max(np.abs(q1.T @ a @ q1 - band)) < (1e-13 if prec == "d" else 3e-6) * n


@pytest.mark.parametrize("prec,b,n", [("d", 32, 260), ("d", 8, 131), ("s", 16, 200), ("d", 3, 40), ("d", 16, 300),
                                      ("s", 8, 517), ("s", 32, 100),
                                      # the lane-per-column chase at a few bulges per sweep and at b = n - 1
                                      ("d", 8, 9), ("s", 8, 11), ("d", 16, 17), ("d", 16, 40), ("s", 16, 33),
                                      # the bulge back-transform's LDS strip with 2 wavefronts, and past its LDS
                                      ("d", 16, 700), ("d", 16, 1600),
                                      # band widths past 32: the scalar chase and the 64-wide bulge window
                                      ("d", 64, 300), ("s", 48, 250), ("d", 48, 600), ("s", 64, 65),
                                      # the lane chase with its band split between device memory and LDS
                                      ("d", 8, 1300), ("d", 8, 2000), ("s", 16, 1300),
                                      # b = 32: one wavefront per bulge, 16 rows per lane (whole band in LDS, split)
                                      ("d", 32, 33), ("s", 32, 47), ("d", 32, 700), ("s", 32, 1300), ("d", 32, 1788)])
def test_chase(handle, torch_cuda, prec, b, n):
    torch = torch_cuda
    rng = np.random.default_rng(5)
    band = rng.standard_normal((n, n))
    band = np.tril(band + band.T)
    band = np.triu(band, -b)
    full = band + np.tril(band, -1).T
    ld = 2 * b - 1
    ab = np.zeros(xtbeig.band_elems(n, b))
    for j in range(n):
        for dd in range(min(b + 1, n - j)):
            ab[j + dd + ld * j] = full[j + dd, j]
    dt = torch.float64 if prec == "d" else torch.float32
    AB = torch.from_numpy(ab).cuda().to(dt)
    V2 = torch.zeros((n, n), dtype=dt, device="cuda")
    D = torch.zeros(n, dtype=dt, device="cuda")
    E = torch.zeros(n, dtype=dt, device="cuda")
    handle.chase(n, b, AB, V2, D, E)
    assert handle.synchronize() == [0, 0, 0, 0]
    d, e = D.double().cpu().numpy(), E.double().cpu().numpy()
    if prec == "d" and n <= (1000 if b < 32 else 300):
        # same reflectors as the oracle's sequential chase: same tridiagonal to rounding (in FP32, and over the
        # longer sweeps of large n, the tridiagonal form itself drifts with rounding, faster with longer reflectors:
        # at b = 32, n = 700 the oracle's own d moves by 2e-9 at a few positions when its input is perturbed by one
        # ulp, and b = 48, n = 600 gives 1e-10; only the spectrum is compared there)
        od, oe, _ = sbr.chase_bulges(full, b)
        assert np.max(np.abs(d - od)) < 1e-13 * n
        assert np.max(np.abs(np.abs(e[: n - 1]) - np.abs(oe))) < 1e-13 * n
    wt = sl.eigvalsh_tridiagonal(d, e[: n - 1])
    wref = np.linalg.eigvalsh(full)
    tol = 1e-12 if prec == "d" else 2e-4
    assert np.max(np.abs(wt - wref)) < tol * np.max(np.abs(wref))
    # back-transform: eigenvectors of the band matrix from those of T
    wtt, zt = sl.eigh_tridiagonal(d, e[: n - 1])
    Z = torch.from_numpy(np.ascontiguousarray(zt.T)).cuda().to(dt)
    handle.back_bulges(n, b, V2, Z, n)
    handle.synchronize()
    z = host(Z).astype(np.float64)
    r = full @ z - z * wtt
    tol = 1e-12 if prec == "d" else 5e-5
    assert np.linalg.norm(r) / (n * np.linalg.norm(full, 2)) < tol


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 64, 131, 412, 1204])
def test_stedc(handle, torch_cuda, n):
    torch = torch_cuda
    rng = np.random.default_rng(100 + n)
    d = rng.standard_normal(n)
    e = rng.standard_normal(max(n - 1, 0))
    D = torch.from_numpy(d.copy()).cuda()
    E = torch.from_numpy(np.concatenate([e, [0.0]])).cuda()
    Z = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    handle.dstedc(D, E, Z)
    assert handle.synchronize() == [0, 0, 0, 0]
    w, z = D.cpu().numpy(), host(Z)
    wref = sl.eigvalsh_tridiagonal(d, e) if n > 1 else d
    scale = max(np.max(np.abs(d)), np.max(np.abs(e)) if n > 1 else 0)
    assert np.all(np.diff(w) >= 0)
    assert np.max(np.abs(w - wref)) < 1e-13 * n * scale + 1e-300
    t = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert np.linalg.norm(t @ z - z * w) / (n * scale) < 1e-14
    assert np.max(np.abs(z.T @ z - np.eye(n))) < 1e-13 * n


def test_stedc_degenerate(handle, torch_cuda):
    """Deflation paths: a diagonal (every pole deflates), a matrix with repeated diagonal blocks, and
    Wilkinson W21+ (close eigenvalue pairs)."""
    torch = torch_cuda
    cases = []
    cases.append((np.arange(40, dtype=float)[::-1].copy(), np.zeros(39)))
    d = np.tile([1.0, 2.0, 3.0], 20)
    e = np.tile([0.5, 0.5, 0.0], 20)[:-1]
    cases.append((d, e))
    m = 10
    cases.append((np.abs(np.arange(-m, m + 1, dtype=float)), np.ones(2 * m)))
    for d, e in cases:
        n = d.size
        D = torch.from_numpy(d.copy()).cuda()
        E = torch.from_numpy(np.concatenate([e, [0.0]])).cuda()
        Z = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        handle.dstedc(D, E, Z)
        assert handle.synchronize() == [0, 0, 0, 0]
        w, z = D.cpu().numpy(), host(Z)
        t = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        wref = np.linalg.eigvalsh(t)
        assert np.max(np.abs(w - wref)) < 1e-13 * n
        assert np.linalg.norm(t @ z - z * w) < 1e-13 * n
        assert np.max(np.abs(z.T @ z - np.eye(n))) < 1e-13 * n


@pytest.mark.parametrize("n", [5, 17, 64, 131, 412])
@pytest.mark.parametrize("prec", ["d", "s"])
def test_syevd(handle, torch_cuda, n, prec):
    torch = torch_cuda
    a, lam = workloads.standard_matrix(n, 1000)
    dt = torch.float64 if prec == "d" else torch.float32
    A = dev(torch, a, dt)
    W = torch.zeros(n, dtype=torch.float64, device="cuda")
    Z = torch.zeros((n, n), dtype=dt, device="cuda")
    if prec == "d":
        handle.dsyevd(A, W, Z)
    else:
        handle.ssyevd(A, W, Z)
    assert handle.synchronize() == [0, 0, 0, 0]
    w, z = W.cpu().numpy(), host(Z).astype(np.float64)
    g = workloads.gates(a, None, w, z, lam)
    if prec == "d":
        assert workloads.passes(g), g
    else:
        assert g["eig_err"] < 1e-5 and g["orth"] < 1e-5 * n, g


def test_device_calls_refuse_mistyped_buffers(handle, torch_cuda):
    # the FP32 standard solve writes FP64 eigenvalues: an FP32 w (half the bytes) is refused before any launch, as
    # are host tensors and buffers too small for the result
    torch = torch_cuda
    n = 64
    A = torch.eye(n, dtype=torch.float32, device="cuda")
    Z = torch.zeros((n, n), dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError, match="float64"):
        handle.ssyevd(A, torch.zeros(n, dtype=torch.float32, device="cuda"), Z)
    with pytest.raises(ValueError, match="device tensor"):
        handle.ssyevd(A, torch.zeros(n, dtype=torch.float64), Z)
    with pytest.raises(ValueError, match="at least"):
        handle.dsyevd(A.double(), torch.zeros(n - 1, dtype=torch.float64, device="cuda"), Z.double())
    W = torch.zeros(n, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="float64"):
        handle.sygvd(A, A, W, Z, xtbeig.FP64)
    assert handle.synchronize() == [0, 0, 0, 0]
    handle.ssyevd(A, W, Z)
    assert handle.synchronize() == [0, 0, 0, 0]
    assert float((W - 1).abs().max()) < 1e-6


@pytest.mark.parametrize("n", [17, 131, 412])
@pytest.mark.parametrize("mixed", [False, True])
def test_sygvd_device(handle, torch_cuda, n, mixed):
    torch = torch_cuda
    h, s = workloads.generalized_pair(n, 1001)
    wref = sl.eigh(h, s, eigvals_only=True)
    H, S = dev(torch, h), dev(torch, s)
    W = torch.zeros(n, dtype=torch.float64, device="cuda")
    C = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    handle.sygvd(H, S, W, C, xtbeig.FP32_REFINE if mixed else xtbeig.FP64)
    assert handle.synchronize() == [0, 0, 0, 0]
    g = workloads.gates(h, s, W.cpu().numpy(), host(C), wref)
    assert workloads.passes(g), g
