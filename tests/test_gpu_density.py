"""Density matrix P = C f(eps) C^T of the generalized eigensystem and its backward (xtbeig_dev_density,
xtbeig_dev_density_backward[_recompute]) on the GPU.

Pins (no reference implementation exists for the backward):
  * forward: numpy's P from scipy's eigh(H, S) with mu bisected on the host;
  * backward, well-separated spectrum: torch autograd on the CPU through cholesky + eigh (an independent route to
    the same derivative: torch differentiates eigh's eigenvectors, the kernel uses Loewner matrices);
  * backward, the density configuration's spectrum with near-degenerate pairs (gaps down to 1e-10): central finite
    differences of the host forward with step 1e-6 (five-point stencil), to 1e-7 relative;
  * the torch.autograd.Function wrapper, the batched call and the recompute variant against each other.
"""
import numpy as np
import pytest
import scipy.linalg as sl
from scipy.special import expit  # 1 / (1 + e^-x) without overflow

import workloads
import xtbeig

pytestmark = pytest.mark.gpu

KTS = (9.5e-4, 3.17e-3)


def np_mu(w, kT, nel):
    lo, hi = w.min() - 60 * kT - 1, w.max() + 60 * kT + 1
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        cnt = np.sum(expit(-(w - mid) / kT))
        lo, hi = (mid, hi) if cnt < nel else (lo, mid)
    return 0.5 * (lo + hi)


def np_density(h, s, kT, nel, mu=None):
    w, c = sl.eigh(h, s)
    if mu is None:
        mu = np_mu(w, kT, nel)
    f = expit(-(w - mu) / kT)
    return (c * f) @ c.T, mu, w, c


def to_dev(torch, mats):
    """(batch, n, n) device tensor whose [p] is mats[p]^T in C order = mats[p] column-major."""
    return torch.from_numpy(np.ascontiguousarray(np.stack([m.T for m in mats]))).cuda()


def from_dev(t):
    return np.stack([m.T for m in t.cpu().numpy()])


def random_sym(n, rng):
    g = rng.standard_normal((n, n))
    return 0.5 * (g + g.T)


def separated_problem(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    s = np.eye(n) + 0.1 * (a @ a.T) / n
    lam = np.sort(rng.uniform(-0.9, 0.6, n))
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    u = np.linalg.cholesky(s).T
    h = u.T @ ((q * lam) @ q.T) @ u
    return 0.5 * (h + h.T), s


@pytest.mark.parametrize("kT", KTS)
def test_density_forward(handle, torch_cuda, kT):
    torch = torch_cuda
    n = 64
    h, s, _ = workloads.density_problem(n, workloads.SEED_BASE + 4)
    pn, mun, w, c = np_density(h, s, kT, n / 2)
    wd = torch.from_numpy(w[None, :].copy()).cuda()
    cd = to_dev(torch, [c])
    mu = torch.zeros(1, dtype=torch.float64, device="cuda")
    p, occ = handle.density(wd, cd, kT, n / 2, mu)
    handle.synchronize()
    assert abs(mu.item() - mun) < 1e-12
    pg = from_dev(p)[0]
    assert np.max(np.abs(pg - pn)) < 1e-12
    assert abs(np.trace(pg @ s) - n / 2) < 1e-10


def torch_cpu_reference(h, s, kT, nel, gbar, mu_fixed):
    """dL/dH, dL/dS of L = <gbar, P(H, S)> by torch autograd on the CPU (cholesky + eigh), symmetrized."""
    import torch
    H = torch.tensor(h, dtype=torch.float64, requires_grad=True)
    S = torch.tensor(s, dtype=torch.float64, requires_grad=True)
    L = torch.linalg.cholesky(S)
    Li = torch.linalg.inv(L)
    hp = Li @ H @ Li.T
    e, Q = torch.linalg.eigh(0.5 * (hp + hp.T))
    C = Li.T @ Q
    if mu_fixed is not None:
        mu = torch.tensor(mu_fixed, dtype=torch.float64)
    else:
        mu0 = np_mu(e.detach().numpy(), kT, nel)
        f0 = torch.sigmoid(-(e - mu0) / kT)
        fp = -f0 * (1 - f0) / kT
        mu = mu0 + (f0.sum() - nel) / (-fp.sum()) * (-1.0)  # one Newton step: value mu0, implicit-function gradient
    f = torch.sigmoid(-(e - mu) / kT)
    P = (C * f) @ C.T
    loss = (torch.tensor(gbar) * P).sum()
    loss.backward()
    gh, gs = H.grad.numpy(), S.grad.numpy()
    return 0.5 * (gh + gh.T), 0.5 * (gs + gs.T)


@pytest.mark.parametrize("mu_fixed", [True, False])
@pytest.mark.parametrize("kT", KTS)
def test_backward_matches_torch_autograd(handle, torch_cuda, kT, mu_fixed):
    torch = torch_cuda
    n = 48
    h, s = separated_problem(n, 77)
    rng = np.random.default_rng(5)
    gbar = random_sym(n, rng)
    _, mun, w, c = np_density(h, s, kT, n / 2)
    rh, rs = torch_cpu_reference(h, s, kT, n / 2, gbar, mun if mu_fixed else None)
    wd = torch.from_numpy(w[None, :].copy()).cuda()
    cd = to_dev(torch, [c])
    mu = torch.zeros(1, dtype=torch.float64, device="cuda")
    _, occ = handle.density(wd, cd, kT, n / 2, mu)
    hb, sb = handle.density_backward(wd, cd, kT, mu, occ, to_dev(torch, [gbar]), mu_fixed=mu_fixed)
    handle.synchronize()
    hg, sg = from_dev(hb)[0], from_dev(sb)[0]
    assert np.max(np.abs(hg - rh)) < 1e-8 * max(1.0, np.max(np.abs(rh)))
    assert np.max(np.abs(sg - rs)) < 1e-8 * max(1.0, np.max(np.abs(rs)))


def fd5(fun, h):
    return (-fun(2 * h) + 8 * fun(h) - 8 * fun(-h) + fun(-2 * h)) / (12 * h)


@pytest.mark.parametrize("kT", KTS)
def test_backward_matches_finite_differences(handle, torch_cuda, kT):
    """The density configuration's spectrum (8 near-degenerate pairs, gaps 1e-10 .. 1e-4), mu re-bisected."""
    torch = torch_cuda
    n = 96
    h, s, _ = workloads.density_problem(n, workloads.SEED_BASE + 4)
    rng = np.random.default_rng(11)
    gbar = random_sym(n, rng)
    dh = random_sym(n, rng)
    dh /= np.linalg.norm(dh)
    ds = random_sym(n, rng)
    ds /= np.linalg.norm(ds)
    hd, sd = to_dev(torch, [h]), to_dev(torch, [s])
    gd = to_dev(torch, [gbar])
    hb, sb = handle.density_backward_recompute(hd, sd, kT, n / 2, gd)
    handle.synchronize()
    hg, sg = from_dev(hb)[0], from_dev(sb)[0]

    def loss(hh, ss):
        p, _, _, _ = np_density(hh, ss, kT, n / 2)
        return np.sum(gbar * p)

    # along the gradient itself (directional derivative = |grad|): 1e-7 relative
    uh, us = hg / np.linalg.norm(hg), sg / np.linalg.norm(sg)
    fh = fd5(lambda t: loss(h + t * uh, s), 1e-6)
    fs = fd5(lambda t: loss(h, s + t * us), 1e-6)
    assert abs(np.linalg.norm(hg) - fh) <= 1e-7 * abs(fh)
    assert abs(np.linalg.norm(sg) - fs) <= 1e-7 * abs(fs)
    # along random unit directions: error within 1e-7 |grad| (the derivative itself can be small there)
    fh = fd5(lambda t: loss(h + t * dh, s), 1e-6)
    fs = fd5(lambda t: loss(h, s + t * ds), 1e-6)
    assert abs(np.sum(hg * dh) - fh) <= 1e-7 * np.linalg.norm(hg)
    assert abs(np.sum(sg * ds) - fs) <= 1e-7 * np.linalg.norm(sg)


def test_batched_equals_single_and_recompute(handle, torch_cuda):
    torch = torch_cuda
    n, batch, kT = 64, 4, 9.5e-4
    probs = [workloads.density_problem(n, 2000 + q)[:2] for q in range(batch)]
    rng = np.random.default_rng(3)
    gb = [random_sym(n, rng) for _ in range(batch)]
    ws, cs = [], []
    for h, s in probs:
        w, c = sl.eigh(h, s)
        ws.append(w)
        cs.append(c)
    wd = torch.from_numpy(np.stack(ws)).cuda()
    cd = to_dev(torch, cs)
    mu = torch.zeros(batch, dtype=torch.float64, device="cuda")
    _, occ = handle.density(wd, cd, kT, n / 2, mu)
    hb, sb = handle.density_backward(wd, cd, kT, mu, occ, to_dev(torch, gb))
    hr, sr = handle.density_backward_recompute(to_dev(torch, [p[0] for p in probs]), to_dev(torch, [p[1] for p in probs]),
                                               kT, n / 2, to_dev(torch, gb))
    handle.synchronize()
    for q in range(batch):
        w1 = wd[q:q + 1].contiguous()
        c1 = cd[q:q + 1].contiguous()
        m1 = torch.zeros(1, dtype=torch.float64, device="cuda")
        _, o1 = handle.density(w1, c1, kT, n / 2, m1)
        h1, s1 = handle.density_backward(w1, c1, kT, m1, o1, to_dev(torch, [gb[q]]))
        handle.synchronize()
        assert torch.equal(h1[0], hb[q]) and torch.equal(s1[0], sb[q])
        # recomputed eigenvectors differ from scipy's by rounding and sign; the gradient does not depend on either
        scale = torch.max(torch.abs(hb[q])).item()
        assert torch.max(torch.abs(hr[q] - hb[q])).item() < 1e-9 * scale
        assert torch.max(torch.abs(sr[q] - sb[q])).item() < 1e-9 * max(scale, 1.0)


class GpuDensity:
    """torch.autograd.Function over the GPU path: forward = generalized eigensolve + density matrix; backward = the
    Loewner-matrix kernel on stored eigenvectors, or on recomputed ones (recompute=True stores only H and S)."""

    @staticmethod
    def make(handle, kT, nel, recompute):
        import torch

        class F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, H, S):
                n = H.shape[0]
                w = torch.zeros((1, n), dtype=torch.float64, device="cuda")
                c = torch.zeros((1, n, n), dtype=torch.float64, device="cuda")
                hT, sT = H.T.contiguous(), S.T.contiguous()
                handle.sygvd(hT, sT, w[0], c[0], xtbeig.FP64)
                mu = torch.zeros(1, dtype=torch.float64, device="cuda")
                p, occ = handle.density(w, c, kT, nel, mu)
                handle.synchronize()
                if recompute:
                    ctx.save_for_backward(hT, sT)
                else:
                    ctx.save_for_backward(w, c, mu, occ)
                return p[0].T.contiguous()

            @staticmethod
            def backward(ctx, g):
                gT = g.T.contiguous()[None]
                if recompute:
                    hT, sT = ctx.saved_tensors
                    hb, sb = handle.density_backward_recompute(hT[None], sT[None], kT, nel, gT)
                else:
                    w, c, mu, occ = ctx.saved_tensors
                    hb, sb = handle.density_backward(w, c, kT, mu, occ, gT)
                handle.synchronize()
                return hb[0].T.contiguous(), sb[0].T.contiguous()

        return F


@pytest.mark.parametrize("recompute", [False, True])
def test_autograd_function(handle, torch_cuda, recompute):
    torch = torch_cuda
    n, kT = 48, 3.17e-3
    h, s = separated_problem(n, 91)
    rng = np.random.default_rng(8)
    gbar = random_sym(n, rng)
    rh, rs = torch_cpu_reference(h, s, kT, n / 2, gbar, None)
    F = GpuDensity.make(handle, kT, n / 2, recompute)
    H = torch.tensor(h, device="cuda", requires_grad=True)
    S = torch.tensor(s, device="cuda", requires_grad=True)
    P = F.apply(H, S)
    (P * torch.tensor(gbar, device="cuda")).sum().backward()
    assert np.max(np.abs(H.grad.cpu().numpy() - rh)) < 1e-8 * max(1.0, np.max(np.abs(rh)))
    assert np.max(np.abs(S.grad.cpu().numpy() - rs)) < 1e-8 * max(1.0, np.max(np.abs(rs)))


def np_density_spin(h, s, kT, na, nb):
    w, c = sl.eigh(h, s)
    ma, mb = np_mu(w, kT, na), np_mu(w, kT, nb)
    fa = expit(-(w - ma) / kT)
    fb = expit(-(w - mb) / kT)
    return (c * (fa + fb)) @ c.T, (ma, mb), w, c, fa, fb


@pytest.mark.parametrize("kT", KTS)
def test_density_spin_forward_and_closed_shell_limit(handle, torch_cuda, kT):
    """Open-shell occupations (the reference's alpha / beta smearing, refine.f90:85-93) against numpy; with equal
    counts the open-shell P and gradient are exactly twice the closed-shell ones of the same count."""
    torch = torch_cuda
    n = 96
    h, s, _ = workloads.density_problem(n, workloads.SEED_BASE + 4)
    na, nb = n / 2 + 2, n / 2 - 2
    pn, (ma, mb), w, c, fa, fb = np_density_spin(h, s, kT, na, nb)
    wd = torch.from_numpy(w[None, :].copy()).cuda()
    cd = to_dev(torch, [c])
    mu = torch.zeros((1, 2), dtype=torch.float64, device="cuda")
    p, occ = handle.density_spin(wd, cd, kT, na, nb, mu)
    handle.synchronize()
    # mu itself is not pinned where a count falls in a gap wider than kT (the count is flat there to rounding): the
    # counts it gives and P are
    for q, nq in ((0, na), (1, nb)):
        cnt = np.sum(expit(-(w - mu[0, q].item()) / kT))
        assert abs(cnt - nq) < 1e-10, (q, cnt, nq)
    pg = from_dev(p)[0]
    assert np.max(np.abs(pg - pn)) < 1e-12
    assert abs(np.trace(pg @ s) - (na + nb)) < 1e-10
    o = occ.cpu().numpy()[0]
    assert abs(o[:n].sum() - na) < 1e-10 and abs(o[3 * n:4 * n].sum() - nb) < 1e-10
    # equal counts: exactly twice the closed-shell result
    m = n / 2
    mu2 = torch.zeros((1, 2), dtype=torch.float64, device="cuda")
    p2, occ2 = handle.density_spin(wd, cd, kT, m, m, mu2)
    mu1 = torch.zeros(1, dtype=torch.float64, device="cuda")
    p1, occ1 = handle.density(wd, cd, kT, m, mu1)
    gb = to_dev(torch, [random_sym(n, np.random.default_rng(4))])
    hb2, sb2 = handle.density_backward_spin(wd, cd, kT, mu2, occ2, gb)
    hb1, sb1 = handle.density_backward(wd, cd, kT, mu1, occ1, gb)
    handle.synchronize()
    assert torch.equal(p2, 2 * p1)
    assert torch.equal(hb2, 2 * hb1) and torch.equal(sb2, 2 * sb1)


@pytest.mark.parametrize("kT", KTS)
def test_density_spin_backward_finite_differences(handle, torch_cuda, kT):
    """Open-shell backward (both chemical potentials re-bisected) against five-point finite differences at h = 1e-6,
    1e-7 relative, on the spectrum with near-degenerate pairs."""
    torch = torch_cuda
    n = 96
    h, s, _ = workloads.density_problem(n, workloads.SEED_BASE + 4)
    na, nb = n / 2 + 1, n / 2 - 1
    rng = np.random.default_rng(21)
    gbar = random_sym(n, rng)
    _, _, w, c, _, _ = np_density_spin(h, s, kT, na, nb)
    wd = torch.from_numpy(w[None, :].copy()).cuda()
    cd = to_dev(torch, [c])
    mu = torch.zeros((1, 2), dtype=torch.float64, device="cuda")
    _, occ = handle.density_spin(wd, cd, kT, na, nb, mu)
    hb, sb = handle.density_backward_spin(wd, cd, kT, mu, occ, to_dev(torch, [gbar]))
    handle.synchronize()
    hg, sg = from_dev(hb)[0], from_dev(sb)[0]

    def loss(hh, ss):
        return np.sum(gbar * np_density_spin(hh, ss, kT, na, nb)[0])

    uh, us = hg / np.linalg.norm(hg), sg / np.linalg.norm(sg)
    fh = fd5(lambda t: loss(h + t * uh, s), 1e-6)
    fs = fd5(lambda t: loss(h, s + t * us), 1e-6)
    assert abs(np.linalg.norm(hg) - fh) <= 1e-7 * abs(fh)
    assert abs(np.linalg.norm(sg) - fs) <= 1e-7 * abs(fs)
