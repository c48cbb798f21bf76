"""Synthetic eigenproblems of the benchmark configurations and the accuracy gates.

Inputs (all column-major float64, symmetric):
  * spectrum(n): 45 % uniform in [-0.9, -0.25], 10 % in [-0.25, -0.15], the rest in [0.0, 0.6]; then 6 exactly
    repeated pairs and 4 pairs split by 1e-9 (entries overwritten in place).
  * standard_matrix(n): H = Q diag(lambda) Q^T with Q = H_1 ... H_n, H_i = I - 2 v v^T / (v^T v), v ~ N(0, 1)^n.
  * overlap_matrix(n): S = I/2 + G/2, G_ij = exp(-|r_i - r_j|^2 / (2 * 1.2^2)) for n points uniform in a cube of
    side (n / 0.8)^(1/3). (The construction gives cond(S) of about 20 at these n: the I/2 shift bounds the smallest
    eigenvalue by 1/2.)
  * overlap_matrix_cond(n, cond): the same G shifted and scaled to cond_2(S) = cond (1e3 - 1e4 in the
    configurations); generalized_pair_cond pairs it with H either overlap-weighted (the generalized spectrum is the
    spectrum above) or independent of S as in generalized_pair.
Seeds are fixed per configuration: 1000 + configuration index.

Gates (the reference's accuracy quantities, Sources/EigenBench/main.swift:233-251 and scripts/kernels/
eigbench.f90:189-204, with the thresholds of the generalized-solve contract):
  * eigenvalue error  max |w - w_ref| <= 1e-12 ||H||_2
  * residual          ||H C - S C diag(w)||_F / (n ||H||_2) <= 1e-13
  * orthogonality     max |C^T S C - I| <= 1e-12 n
"""
import numpy as np

SEED_BASE = 1000


def spectrum(n, rng):
    n1 = int(round(0.45 * n))
    n2 = int(round(0.10 * n))
    lam = np.concatenate([
        rng.uniform(-0.9, -0.25, n1),
        rng.uniform(-0.25, -0.15, n2),
        rng.uniform(0.0, 0.6, n - n1 - n2),
    ])
    if n >= 20:
        pick = rng.permutation(n)[:20]
        for p in range(6):
            lam[pick[2 * p + 1]] = lam[pick[2 * p]]
        for p in range(6, 10):
            lam[pick[2 * p + 1]] = lam[pick[2 * p]] + 1e-9
    return lam


def householder_product(n, rng, block=32):
    """Q = H_1 H_2 ... H_n with H_i = I - 2 v_i v_i^T / (v_i^T v_i), v_i ~ N(0, 1)^n, applied block-wise
    (compact WY, Q = I - Y T Y^T per block of reflectors) so that it costs GEMMs."""
    q = np.eye(n)
    i = 0
    while i < n:
        k = min(block, n - i)
        v = rng.standard_normal((k, n)).T  # one N(0, 1)^n vector per reflector, in order
        tau = 2.0 / np.einsum("ij,ij->j", v, v)
        # T for H_i ... H_{i+k-1} = I - V T V^T (forward, columnwise)
        t = np.zeros((k, k))
        g = v.T @ v
        for j in range(k):
            t[j, j] = tau[j]
            if j:
                t[:j, j] = -tau[j] * (t[:j, :j] @ g[:j, j])
        # Q <- Q (I - V T V^T)
        q = q - (q @ v) @ t @ v.T
        i += k
    return q


def standard_matrix(n, seed):
    rng = np.random.default_rng(seed)
    lam = spectrum(n, rng)
    q = householder_product(n, rng)
    h = (q * lam) @ q.T
    h = 0.5 * (h + h.T)
    return np.asfortranarray(h), np.sort(lam)


def overlap_matrix(n, rng):
    side = (n / 0.8) ** (1.0 / 3.0)
    r = rng.uniform(0.0, side, (n, 3))
    sq = np.einsum("ij,ij->i", r, r)
    d2 = np.maximum(sq[:, None] + sq[None, :] - 2.0 * (r @ r.T), 0.0)
    s = 0.5 * np.eye(n) + 0.5 * np.exp(-d2 / (2.0 * 1.2 ** 2))
    return np.asfortranarray(0.5 * (s + s.T))


def gaussian_overlap(n, rng):
    """G_ij = exp(-|r_i - r_j|^2 / (2 * 1.2^2)) for n points uniform in a cube of side (n / 0.8)^(1/3)."""
    side = (n / 0.8) ** (1.0 / 3.0)
    r = rng.uniform(0.0, side, (n, 3))
    sq = np.einsum("ij,ij->i", r, r)
    d2 = np.maximum(sq[:, None] + sq[None, :] - 2.0 * (r @ r.T), 0.0)
    g = np.exp(-d2 / (2.0 * 1.2 ** 2))
    return 0.5 * (g + g.T)


def overlap_matrix_cond(n, rng, cond):
    """S = (G + delta I) / (1 + delta), G the Gaussian overlap, delta chosen so that cond_2(S) = cond exactly
    (delta = (g_max - cond g_min) / (cond - 1) from G's extreme eigenvalues); unit diagonal like an overlap matrix.
    The configurations ask for cond(S) of 1e3 - 1e4, which S = I/2 + G/2 cannot reach."""
    g = gaussian_overlap(n, rng)
    ev = np.linalg.eigvalsh(g)
    delta = (ev[-1] - cond * ev[0]) / (cond - 1.0)
    s = (g + delta * np.eye(n)) / (1.0 + delta)
    return np.asfortranarray(0.5 * (s + s.T))


def generalized_pair_cond(n, seed, cond, physical=True):
    """(H, S, lambda) with cond_2(S) = cond. physical: H = U^T (Q diag(lambda) Q^T) U with S = U^T U, so the
    generalized eigenvalues are exactly the spectrum lambda (as for an overlap-weighted Hamiltonian, and what the
    degenerate and near-degenerate pairs of the spectrum test); otherwise H = Q diag(lambda) Q^T as in
    generalized_pair, whose generalized eigenvalues grow to about ||H|| / lambda_min(S) (lambda is then not the
    generalized spectrum and None is returned for it)."""
    rng = np.random.default_rng(seed)
    lam = spectrum(n, rng)
    q = householder_product(n, rng)
    m = (q * lam) @ q.T
    m = 0.5 * (m + m.T)
    s = overlap_matrix_cond(n, rng, cond)
    if physical:
        u = np.linalg.cholesky(s).T
        h = u.T @ m @ u
        return np.asfortranarray(0.5 * (h + h.T)), s, np.sort(lam)
    return np.asfortranarray(m), s, None


def generalized_pair(n, seed):
    """(H, S) of the generalized configurations: H as standard_matrix, S = overlap_matrix, one seed."""
    rng = np.random.default_rng(seed)
    lam = spectrum(n, rng)
    q = householder_product(n, rng)
    h = (q * lam) @ q.T
    h = np.asfortranarray(0.5 * (h + h.T))
    s = overlap_matrix(n, rng)
    return h, s


def density_problem(n, seed):
    """(H, S, lambda) of the density-matrix configuration: the generalized eigenvalues are the spectrum above plus 8
    near-degenerate pairs with gaps log-uniform in [1e-10, 1e-4]. H = U^T (Q diag(lambda) Q^T) U with S = U^T U, so
    (H, S) has exactly the spectrum lambda."""
    rng = np.random.default_rng(seed)
    lam = spectrum(n, rng)
    if n >= 40:
        pick = rng.permutation(n)[:16]
        gaps = 10.0 ** rng.uniform(-10.0, -4.0, 8)
        for p in range(8):
            lam[pick[2 * p + 1]] = lam[pick[2 * p]] + gaps[p]
    q = householder_product(n, rng)
    s = overlap_matrix(n, rng)
    u = np.linalg.cholesky(s).T
    hp = (q * lam) @ q.T
    h = u.T @ (0.5 * (hp + hp.T)) @ u
    return np.asfortranarray(0.5 * (h + h.T)), s, np.sort(lam)


def gates(h, s, w, c, w_ref, scaled=False):
    """The three contract quantities, normalized so that each passes at <= 1.

    scaled: the normalizations of a problem whose generalized eigenvalues are much larger than ||H||_2 (H independent
    of an ill-conditioned S): the eigenvalue gate relative to max(||H||_2, max |w_ref|) and the residual gate relative
    to ||H||_2 + max |w_ref| ||S||_2, the scales of the rounding errors of any FP64 method on such a problem (the FP64
    host LAPACK itself misses the unscaled eigenvalue gate there)."""
    n = h.shape[0]
    nh = np.linalg.norm(h, 2)
    wmax = float(np.max(np.abs(w_ref)))
    eig = np.max(np.abs(np.asarray(w) - np.asarray(w_ref))) / (max(nh, wmax) if scaled else nh)
    if scaled and s is not None:
        nh = nh + wmax * np.linalg.norm(s, 2)
    if s is None:
        r = h @ c - c * w
        o = np.max(np.abs(c.T @ c - np.eye(n)))
    else:
        r = h @ c - (s @ c) * w
        o = np.max(np.abs(c.T @ s @ c - np.eye(n)))
    res = np.linalg.norm(r) / (n * nh)
    return {
        "eig_err": float(eig), "eig_gate": float(eig / 1e-12),
        "residual": float(res), "residual_gate": float(res / 1e-13),
        "orth": float(o), "orth_gate": float(o / (1e-12 * n)),
    }


def passes(g):
    return g["eig_gate"] <= 1 and g["residual_gate"] <= 1 and g["orth_gate"] <= 1
