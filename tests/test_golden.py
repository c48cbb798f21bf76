"""Golden fixtures (tests/golden, written by tests/golden/make_golden.py with the host LAPACK) on the CPU: the HS.bin
reader/writer round trip, and the oracle pinned against the stored vectors (its C restatement of the 1-stage chain
reproduces the stored eigenpairs)."""
import json
import os

import numpy as np
import pytest

import hsbin
import host_lapack as hl
import sbr
import workloads

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SIZES = (5, 17, 64, 131)


def test_hsbin_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    h, s = rng.standard_normal((7, 7)), rng.standard_normal((7, 7))
    p = tmp_path / "x.bin"
    hsbin.write(p, h, s)
    h2, s2, e2, c2 = hsbin.read(p)
    assert np.array_equal(h, h2) and np.array_equal(s, s2) and e2 is None and c2 is None
    e, c = rng.standard_normal(7), rng.standard_normal((7, 7))
    hsbin.write(p, h, s, e, c)
    h2, s2, e2, c2 = hsbin.read(p)
    assert np.array_equal(e, e2) and np.array_equal(c, c2)
    raw = open(p, "rb").read()
    assert np.frombuffer(raw[:4], dtype=np.int32)[0] == 7 and len(raw) == 4 + 8 * (3 * 49 + 7)
    open(p, "wb").write(raw[:-8])
    with pytest.raises(ValueError):
        hsbin.read(p)


def oracle_sygvd(h, s):
    """host dpotrf + dsygst, the C restatement of the 1-stage reduction, host dstedc, the restated dormtr, dtrsm."""
    u, info = hl.dpotrf_upper(s)
    u = np.triu(u)
    hp, info = hl.dsygst_upper(h, u)
    a, d, e, tau = sbr.sytrd_upper(hp)
    w, zt, info = hl.dstedc(d, e)
    z = sbr.ormtr_upper(a, tau, zt)
    return w, hl.dtrsm_left_upper(u, z)


@pytest.mark.parametrize("n", SIZES)
def test_oracle_reproduces_golden(n):
    h, s, e, c = hsbin.read(os.path.join(GOLDEN, f"hs_{n}.bin"))
    w, x = oracle_sygvd(h, s)
    nh = np.linalg.norm(h, 2)
    assert np.max(np.abs(w - e)) <= 1e-13 * nh
    # eigenvectors up to sign, for eigenvalues apart from their neighbours
    gap = np.minimum(np.diff(e, prepend=-np.inf), np.diff(e, append=np.inf))
    sep = gap > 1e-6 * nh
    dots = np.sum(x * (s @ c), axis=0)
    assert np.all(np.abs(np.abs(dots[sep]) - 1) < 1e-10)


def test_c0_spectrum_golden():
    rec = json.load(open(os.path.join(GOLDEN, "c0_412.json")))
    h, _ = workloads.standard_matrix(rec["n"], rec["seed"])
    a, d, e, tau = sbr.sytrd_upper(h)
    w, _, info = hl.dstedc(d, e)
    assert info == 0
    assert np.max(np.abs(w - np.array(rec["eigenvalues"]))) <= 1e-13 * np.linalg.norm(h, 2)
