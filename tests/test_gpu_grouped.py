"""Grouped solves (xtbeig_dev_sygvd_grouped): every kernel of the solve launched once over a group of problems.

Pins:
  * unpadded groups (granule 1): each problem's eigenpairs bit-identical to the eager solve of that problem alone
    (the kernels, their order and every sum are the same; only the grid's z extent and the slices differ);
  * padded groups (granule 64, and 16): n from 1 to 300 padded to the granule; every problem's status words zero,
    its eigenvalues within rounding of the eager solve's, and the three FP64 gates against scipy.linalg.eigh;
  * chunking: a granule that puts every problem of a large count into one group still solves each (the chunks are
    bounded by the launch grid's z limit);
  * a padded block whose diagonal is constant (H' = c I) still separates (the padded values then come from the
    off-diagonal range, or from c itself).
"""
import numpy as np
import pytest
import scipy.linalg as sl

import workloads
import xtbeig

pytestmark = pytest.mark.gpu


def dev(torch, m):
    return torch.from_numpy(np.ascontiguousarray(m.T)).cuda()


def solve_eager(torch, hnd, h, s, prec):
    n = h.shape[0]
    w = torch.zeros(n, dtype=torch.float64, device="cuda")
    c = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    hd, sd = dev(torch, h), dev(torch, s)
    hnd.sygvd(hd, sd, w, c, prec)
    assert hnd.synchronize() == [0, 0, 0, 0]
    return w.cpu().numpy(), c.cpu().numpy()


def grouped(torch, hnd, probs, prec, granule, ld_extra=0):
    hs, ss, ws, cs = [], [], [], []
    for h, s in probs:
        n = h.shape[0]
        ld = n + ld_extra
        hp = torch.zeros((n, ld), dtype=torch.float64, device="cuda")
        sp = torch.zeros((n, ld), dtype=torch.float64, device="cuda")
        hp[:, :n] = dev(torch, h)
        sp[:, :n] = dev(torch, s)
        hs.append(hp)
        ss.append(sp)
        ws.append(torch.full((n,), np.nan, dtype=torch.float64, device="cuda"))
        cs.append(torch.full((n, ld), np.nan, dtype=torch.float64, device="cuda"))
    st, lat = hnd.sygvd_grouped(hs, ss, ws, cs, prec, granule=granule)
    out = [(ws[q].cpu().numpy(), cs[q][:, :probs[q][0].shape[0]].cpu().numpy()) for q in range(len(probs))]
    # the padding past ld is never written
    for q, (h, _) in enumerate(probs):
        n = h.shape[0]
        if ld_extra:
            assert torch.isnan(cs[q][:, n:]).all(), q
    return st, lat, out


@pytest.mark.parametrize("prec", [xtbeig.FP64, xtbeig.FP32_REFINE])
def test_grouped_unpadded_bit_identical(torch_cuda, prec):
    torch = torch_cuda
    ns = [200, 200, 200, 131, 131, 200, 64, 1, 2, 3, 3]
    probs = [workloads.generalized_pair(n, 700 + q) for q, n in enumerate(ns)]
    hnd = xtbeig.Handle(0)
    st, lat, out = grouped(torch, hnd, probs, prec, granule=1, ld_extra=3)
    assert hnd.synchronize() == [0, 0, 0, 0]
    assert all(s == [0, 0, 0, 0] for s in st), st
    assert all(x > 0 for x in lat)
    ref = xtbeig.Handle(0)
    for q, (h, s) in enumerate(probs):
        we, ce = solve_eager(torch, ref, h, s, prec)
        w, c = out[q]
        print(q, ns[q], float(np.max(np.abs(w - we))), float(np.max(np.abs(c - ce))), flush=True)
        assert np.array_equal(w, we) and np.array_equal(c, ce), q
    ref.close()
    hnd.close()


@pytest.mark.parametrize("prec", [xtbeig.FP64, xtbeig.FP32_REFINE])
def test_grouped_band32_bit_identical(torch_cuda, prec):
    # band width 32: the lane-per-column chase with one wavefront per bulge, over the group's z axis, with the band
    # wholly in LDS (small n) and split between device memory and LDS (n = 300 in FP64)
    torch = torch_cuda
    ns = [300, 300, 150, 150, 100, 40]
    probs = [workloads.generalized_pair(n, 900 + q) for q, n in enumerate(ns)]
    hnd = xtbeig.Handle(0)
    ref = xtbeig.Handle(0)
    try:
        hnd.set_tuning(32, 2)
        ref.set_tuning(32, 2)
        st, lat, out = grouped(torch, hnd, probs, prec, granule=1)
        assert hnd.synchronize() == [0, 0, 0, 0]
        assert all(s == [0, 0, 0, 0] for s in st), st
        for q, (h, s) in enumerate(probs):
            we, ce = solve_eager(torch, ref, h, s, prec)
            w, c = out[q]
            assert np.array_equal(w, we) and np.array_equal(c, ce), q
            g = workloads.gates(h, s, w, c.T, sl.eigh(h, s, eigvals_only=True))
            assert workloads.passes(g), (q, g)
    finally:
        ref.close()
        hnd.close()


@pytest.mark.parametrize("granule", [64, 16])
@pytest.mark.parametrize("prec", [xtbeig.FP64, xtbeig.FP32_REFINE])
def test_grouped_padded_uneven(torch_cuda, prec, granule):
    torch = torch_cuda
    rng = np.random.default_rng(1003)
    ns = [1, 2, 3, 17, 64, 65] + [int(round(np.exp(x))) for x in rng.uniform(np.log(40), np.log(300), 10)]
    probs = [workloads.generalized_pair(n, 800 + q) for q, n in enumerate(ns)]
    hnd = xtbeig.Handle(0)
    st, lat, out = grouped(torch, hnd, probs, prec, granule=granule, ld_extra=1)
    assert all(s == [0, 0, 0, 0] for s in st), st
    assert hnd.synchronize() == [0, 0, 0, 0]
    ref = xtbeig.Handle(0)
    for q, (h, s) in enumerate(probs):
        w, c = out[q]
        wref = sl.eigh(h, s, eigvals_only=True)
        g = workloads.gates(h, s, w, c.T, wref)
        assert workloads.passes(g), (q, ns[q], g)
        we, _ = solve_eager(torch, ref, h, s, prec)
        assert np.max(np.abs(w - we)) <= 1e-13 * max(1.0, np.linalg.norm(h, 2)), (q, ns[q])
    ref.close()
    hnd.close()


def test_grouped_constant_diagonal(torch_cuda):
    """H = c I + (off-diagonal) and H = c I: the padded diagonal still lies in the problem's spectral range and
    separates."""
    torch = torch_cuda
    n = 40
    rng = np.random.default_rng(5)
    a = rng.standard_normal((n, n)) * 0.01
    a = a + a.T
    np.fill_diagonal(a, 0.0)
    h1 = 0.3 * np.eye(n) + a
    h2 = -0.7 * np.eye(n)
    s = workloads.overlap_matrix(n, np.random.default_rng(6))
    probs = [(np.asfortranarray(h1), s), (np.asfortranarray(h2), np.asfortranarray(np.eye(n)))]
    hnd = xtbeig.Handle(0)
    st, _, out = grouped(torch, hnd, probs, xtbeig.FP64, granule=64)
    assert all(x == [0, 0, 0, 0] for x in st), st
    for q, (h, sq) in enumerate(probs):
        w, c = out[q]
        g = workloads.gates(h, sq, w, c.T, sl.eigh(h, sq, eigvals_only=True))
        assert workloads.passes(g), (q, g)
    hnd.close()


def test_grouped_chunks_and_empty(torch_cuda):
    torch = torch_cuda
    hnd = xtbeig.Handle(0)
    assert hnd.sygvd_grouped([], [], [], [], xtbeig.FP64) == ([], [])
    # 90 problems of n = 700 padded to 1024 (granule 1024): the D&C's widest level at 1024 (512 merges, a z extent
    # of 1024) caps a launch at 63 problems, so the group runs as two chunks
    n, count = 700, 90
    h, s = workloads.generalized_pair(n, 901)
    hd, sd = dev(torch, h), dev(torch, s)
    ws = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(count)]
    cs = [torch.zeros((n, n), dtype=torch.float64, device="cuda") for _ in range(count)]
    st, lat = hnd.sygvd_grouped([hd] * count, [sd] * count, ws, cs, xtbeig.FP64, granule=1024)
    assert all(x == [0, 0, 0, 0] for x in st), st
    wref = sl.eigh(h, s, eigvals_only=True)
    w0 = ws[0].cpu().numpy()
    assert workloads.passes(workloads.gates(h, s, w0, cs[0].cpu().numpy().T, wref))
    for q in (1, count // 2, count - 1):
        assert np.array_equal(ws[q].cpu().numpy(), w0), q
        assert np.array_equal(cs[q].cpu().numpy(), cs[0].cpu().numpy()), q
    hnd.close()
