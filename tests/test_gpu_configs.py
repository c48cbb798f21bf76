"""The benchmark configurations at their stated sizes, every solve gated against the FP64 host LAPACK.

  * configuration 2: the 20-step SCC-like sequence H_k = H_0 + D_k at n = 1788 (||D_k||_F from 1e-1 down to 1e-7,
    FP32 + refinement while ||D_k||_F >= 1e-3, then FP64), band 32 and the built-in band, every step's eigenpairs
    against scipy.linalg.eigh with the three gates;
  * configuration 3: 256 generalized problems, n log-uniform in [124, 1204] (seed 1003), through the stream-batched
    API and through the grouped API (every kernel launched once per group of equal padded size), each once in FP64
    and once in FP32 + refinement, every solve against scipy.linalg.eigh;
  * configuration 4: the density-matrix backward for 16 problems at n = 588 and 16 at n = 1204 (the spectrum with 8
    near-degenerate pairs, gaps 1e-10 .. 1e-4), kT = 9.5e-4 and 3.17e-3 Eh, mu bisected to trace(P S) = n / 2,
    upstream gradient random symmetric N(0, 1): every problem against an FP64 CPU reference gradient built on
    scipy's eigenpairs (the Loewner-matrix closed form with its divided differences evaluated directly, or by a
    fourth-order Taylor expansion where the gap is below kT / 100), and one problem per size against central finite
    differences of the host forward (five-point stencil, h = 1e-6) to 1e-7 relative.
"""
import functools

import numpy as np
import pytest
import scipy.linalg as sl

import scc_schedule
import workloads
import xtbeig

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda()


# ---- configuration 2 ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("band", [32, 64, 0])
def test_config2_sequence_n1788(torch_cuda, band):
    torch = torch_cuda
    n, steps = 1788, 20
    seed = workloads.SEED_BASE + 2
    h0, s = workloads.generalized_pair(n, seed)
    sd = dev(torch, s)
    hnd = xtbeig.Handle(0)
    try:
        hnd.set_tuning(band, 2)
        hnd.set_scc_precision(1e-3, 160, 1)
        hnd.scc_begin()
        w = torch.zeros(n, dtype=torch.float64, device="cuda")
        c = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        used_all = []
        for k, norm, hk in scc_schedule.sequence(h0, steps, seed + 100):
            used = hnd.scc_sygvd(dev(torch, hk), sd, w, c, norm)
            used_all.append(used)
            wref = sl.eigh(hk, s, eigvals_only=True)
            g = workloads.gates(hk, s, w.cpu().numpy(), c.cpu().numpy().T, wref)
            print(f"step {k} |D|={norm:.2e} {'fp32' if used == xtbeig.FP32_REFINE else 'fp64'} {g}", flush=True)
            assert workloads.passes(g), (k, g)
            assert used == (xtbeig.FP32_REFINE if norm >= 1e-3 else xtbeig.FP64), (k, norm, used)
        settings, counters, rc = hnd.get_scc_precision()
        assert rc == 0 and counters["scc_cycles"] == steps and counters["fp64_restart"] == 0
        assert counters["fp32_cycles"] == sum(u == xtbeig.FP32_REFINE for u in used_all) > 0
    finally:
        hnd.close()


# ---- configuration 3 ------------------------------------------------------------------------------------------

@functools.lru_cache(maxsize=1)
def config3_problems():
    rng = np.random.default_rng(workloads.SEED_BASE + 3)
    ns = [int(round(np.exp(x))) for x in rng.uniform(np.log(124), np.log(1204), 256)]
    probs = [workloads.generalized_pair(n, workloads.SEED_BASE + 3 + 7 * (q + 1)) for q, n in enumerate(ns)]
    refs = [sl.eigh(h, s, eigvals_only=True) for h, s in probs]
    return ns, probs, refs


@pytest.mark.parametrize("prec", [xtbeig.FP64, xtbeig.FP32_REFINE])
def test_config3_batched_256(torch_cuda, prec):
    torch = torch_cuda
    ns, probs, refs = config3_problems()
    assert len(ns) == 256 and min(ns) >= 124 and max(ns) <= 1204
    hs = [dev(torch, h) for h, _ in probs]
    ss = [dev(torch, s) for _, s in probs]
    ws = [torch.zeros(n, dtype=torch.float64, device="cuda") for n in ns]
    cs = [torch.zeros((n, n), dtype=torch.float64, device="cuda") for n in ns]
    hnd = xtbeig.Handle(0)
    try:
        lat = hnd.sygvd_batched(hs, ss, ws, cs, prec)
        st = hnd.synchronize()
        assert st == [0, 0, 0, 0], st
        assert len(lat) == 256 and all(x > 0 for x in lat)
        worst = {"eig_gate": 0.0, "residual_gate": 0.0, "orth_gate": 0.0}
        for q, (h, s) in enumerate(probs):
            g = workloads.gates(h, s, ws[q].cpu().numpy(), cs[q].cpu().numpy().T, refs[q])
            assert workloads.passes(g), (q, ns[q], g)
            for k in worst:
                worst[k] = max(worst[k], g[k])
        print("worst gates over 256 problems", worst, flush=True)
    finally:
        hnd.close()


@pytest.mark.parametrize("prec", [xtbeig.FP64, xtbeig.FP32_REFINE])
def test_config3_grouped_256(torch_cuda, prec):
    """The same 256 problems through xtbeig_dev_sygvd_grouped: padded to multiples of 64 (18 groups), every kernel
    launched once per group, groups on four streams; every problem's status words zero and its gates passed."""
    torch = torch_cuda
    ns, probs, refs = config3_problems()
    hs = [dev(torch, h) for h, _ in probs]
    ss = [dev(torch, s) for _, s in probs]
    ws = [torch.zeros(n, dtype=torch.float64, device="cuda") for n in ns]
    cs = [torch.zeros((n, n), dtype=torch.float64, device="cuda") for n in ns]
    hnd = xtbeig.Handle(0)
    try:
        st, lat = hnd.sygvd_grouped(hs, ss, ws, cs, prec, granule=64, slots=4)
        assert all(x == [0, 0, 0, 0] for x in st), [q for q, x in enumerate(st) if x != [0, 0, 0, 0]]
        assert hnd.synchronize() == [0, 0, 0, 0]
        assert len(lat) == 256 and all(x > 0 for x in lat)
        worst = {"eig_gate": 0.0, "residual_gate": 0.0, "orth_gate": 0.0}
        for q, (h, s) in enumerate(probs):
            g = workloads.gates(h, s, ws[q].cpu().numpy(), cs[q].cpu().numpy().T, refs[q])
            assert workloads.passes(g), (q, ns[q], g)
            for k in worst:
                worst[k] = max(worst[k], g[k])
        print("worst gates over 256 grouped problems", worst, flush=True)
    finally:
        hnd.close()


# ---- configuration 4 ------------------------------------------------------------------------------------------

KTS = (9.5e-4, 3.17e-3)


def fermi(x):
    """f = 1 / (1 + e^x) without overflow."""
    return np.where(x > 0, np.exp(-np.abs(x)) / (1.0 + np.exp(-np.abs(x))), 1.0 / (1.0 + np.exp(-np.abs(x))))


def np_mu(w, kT, nel):
    lo, hi = w.min() - 60 * kT - 1, w.max() + 60 * kT + 1
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        cnt = np.sum(fermi((w - mid) / kT))
        lo, hi = (mid, hi) if cnt < nel else (lo, mid)
    return 0.5 * (lo + hi)


def loewner(w, mu, kT):
    """F_ij = (f_i - f_j) / (w_i - w_j), F_ii = f'(w_i): the direct quotient where |w_i - w_j| >= kT / 100, else the
    Taylor expansion about the midpoint, f'(m) + f'''(m) d^2 / 24 (d = w_i - w_j; the next term is below 1e-10 of
    f' there)."""
    x = (w - mu) / kT
    f = fermi(x)
    d = w[:, None] - w[None, :]
    m = 0.5 * (x[:, None] + x[None, :])
    fm = fermi(m)
    fx = -fm * (1.0 - fm)                                   # df/dx at the midpoint
    fxxx = -fm * (1.0 - fm) * (1.0 - 6.0 * fm + 6.0 * fm * fm)
    dx = d / kT
    taylor = (fx + fxxx * dx * dx / 24.0) / kT
    with np.errstate(divide="ignore", invalid="ignore"):
        direct = (f[:, None] - f[None, :]) / d
    F = np.where(np.abs(dx) >= 1e-2, direct, taylor)
    return F, f, -f * (1.0 - f) / kT


def cpu_backward(h, s, kT, nel, gbar):
    """dL/dH, dL/dS of L = <gbar, P>, P = C diag(f) C^T, mu re-bisected, from scipy's FP64 eigenpairs."""
    w, c = sl.eigh(h, s)
    mu = np_mu(w, kT, nel)
    F, f, fp = loewner(w, mu, kT)
    G = f[:, None] + w[None, :] * F                        # Loewner matrix of w f(w): f_i + w_j F_ij
    gam = c.T @ (0.5 * (gbar + gbar.T)) @ c
    kappa = np.sum(fp * np.diag(gam)) / np.sum(fp)         # the chemical potential's response
    hb = c @ (F * gam - kappa * np.diag(fp)) @ c.T
    sb = -c @ (G * gam - kappa * np.diag(w * fp)) @ c.T
    return 0.5 * (hb + hb.T), 0.5 * (sb + sb.T)


def host_loss(h, s, kT, nel, gbar):
    w, c = sl.eigh(h, s)
    mu = np_mu(w, kT, nel)
    return np.sum(gbar * ((c * fermi((w - mu) / kT)) @ c.T))


def fd5(fun, step):
    return (-fun(2 * step) + 8 * fun(step) - 8 * fun(-step) + fun(-2 * step)) / (12 * step)


@functools.lru_cache(maxsize=2)
def config4_problems(n):
    probs = [workloads.density_problem(n, workloads.SEED_BASE + 4 + 17 * q)[:2] for q in range(16)]
    rng = np.random.default_rng(workloads.SEED_BASE + 4 + n)
    gbars = []
    for _ in range(16):
        g = rng.standard_normal((n, n))
        gbars.append(0.5 * (g + g.T))
    return probs, gbars


def to_batch(torch, mats):
    return torch.from_numpy(np.ascontiguousarray(np.stack([m.T for m in mats]))).cuda()


def from_batch(t):
    return np.stack([m.T for m in t.cpu().numpy()])


@pytest.mark.parametrize("kT", KTS)
@pytest.mark.parametrize("n", [588, 1204])
def test_config4_backward_batch16(torch_cuda, n, kT):
    torch = torch_cuda
    probs, gbars = config4_problems(n)
    nel = n / 2
    hnd = xtbeig.Handle(0)
    try:
        w = torch.zeros((16, n), dtype=torch.float64, device="cuda")
        c = torch.zeros((16, n, n), dtype=torch.float64, device="cuda")
        # the inputs stay referenced until the solves are done: the library reads them on its own stream, and torch's
        # allocator would hand a freed block to the next copy at once
        hs = [dev(torch, h) for h, _ in probs]
        ss = [dev(torch, s) for _, s in probs]
        for q in range(16):
            hnd.sygvd(hs[q], ss[q], w[q], c[q], xtbeig.FP64)
        assert hnd.synchronize() == [0, 0, 0, 0]
        del hs, ss
        mu = torch.zeros(16, dtype=torch.float64, device="cuda")
        p, occ = hnd.density(w, c, kT, nel, mu)
        hb, sb = hnd.density_backward(w, c, kT, mu, occ, to_batch(torch, gbars))
        assert hnd.synchronize() == [0, 0, 0, 0]
        hg, sg = from_batch(hb), from_batch(sb)
        pg = from_batch(p)
        for q, (h, s) in enumerate(probs):
            assert abs(np.trace(pg[q] @ s) - nel) < 1e-8 * n, q
            rh, rs = cpu_backward(h, s, kT, nel, gbars[q])
            eh = np.max(np.abs(hg[q] - rh)) / np.max(np.abs(rh))
            es = np.max(np.abs(sg[q] - rs)) / np.max(np.abs(rs))
            assert eh < 1e-8 and es < 1e-8, (q, eh, es)
        # one problem against central finite differences of the host forward, along the gradient's own direction
        h, s = probs[0]
        uh = hg[0] / np.linalg.norm(hg[0])
        us = sg[0] / np.linalg.norm(sg[0])
        fh = fd5(lambda t: host_loss(h + t * uh, s, kT, nel, gbars[0]), 1e-6)
        fs = fd5(lambda t: host_loss(h, s + t * us, kT, nel, gbars[0]), 1e-6)
        print(f"n={n} kT={kT} |Hbar|={np.linalg.norm(hg[0]):.6e} fd={fh:.6e} |Sbar|={np.linalg.norm(sg[0]):.6e} "
              f"fd={fs:.6e}", flush=True)
        assert abs(np.linalg.norm(hg[0]) - fh) <= 1e-7 * abs(fh)
        assert abs(np.linalg.norm(sg[0]) - fs) <= 1e-7 * abs(fs)
        # the recompute variant (eigenvectors solved again, O(n^2) memory) on two problems of the batch
        hr, sr = hnd.density_backward_recompute(to_batch(torch, [p_[0] for p_ in probs[:2]]),
                                                to_batch(torch, [p_[1] for p_ in probs[:2]]), kT, nel,
                                                to_batch(torch, gbars[:2]))
        assert hnd.synchronize() == [0, 0, 0, 0]
        hr, sr = from_batch(hr), from_batch(sr)
        for q in range(2):
            assert np.max(np.abs(hr[q] - hg[q])) < 1e-9 * np.max(np.abs(hg[q])), q
            assert np.max(np.abs(sr[q] - sg[q])) < 1e-9 * np.max(np.abs(sg[q])), q
    finally:
        hnd.close()
