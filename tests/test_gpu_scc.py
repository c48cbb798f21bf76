"""SCC precision schedule (xtbeig_set_scc_precision / xtbeig_scc_begin / xtbeig_dev_scc_sygvd /
xtbeig_get_scc_precision): the contract of xTB's xtb_setSCCPrecision / xtb_getSCCPrecision, restated from the
reference's scripts/test_sccprecision.py:79-161 for the eigensolve level (default FP64; settings reported back;
invalid values rejected with the settings unchanged; the getter refuses before any cycle; two threads with their own
handles each bit-identical to their serial results), and the mixed-precision configuration's SCC-like sequence
checked against the FP64 oracle at every step."""
import threading

import numpy as np
import pytest
import scipy.linalg as sl

import workloads
import xtbeig
from scc_schedule import sequence

pytestmark = pytest.mark.gpu


def setup(torch, n, seed):
    h, s = workloads.generalized_pair(n, seed)
    sd = torch.from_numpy(np.ascontiguousarray(s.T)).cuda()
    w = torch.zeros(n, dtype=torch.float64, device="cuda")
    c = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    return h, s, sd, w, c


def run_sequence(torch, hnd, h0, s, sd, w, c, steps=20, seed=5):
    hnd.scc_begin()
    used, res = [], []
    for k, norm, hk in sequence(h0, steps, seed):
        hd = torch.from_numpy(np.ascontiguousarray(hk.T)).cuda()
        used.append(hnd.scc_sygvd(hd, sd, w, c, norm))
        res.append((w.cpu().numpy().copy(), c.cpu().numpy().copy(), hk))
    return used, res


def test_default_is_fp64_and_getter_needs_a_cycle(torch_cuda):
    torch = torch_cuda
    hnd = xtbeig.Handle(0)
    settings, counters, rc = hnd.get_scc_precision()
    assert rc != 0 and settings == (0.0, 0, 0)
    h, s, sd, w, c = setup(torch, 64, 7)
    used, _ = run_sequence(torch, hnd, h, s, sd, w, c, steps=4)
    assert used == [xtbeig.FP64] * 4
    settings, counters, rc = hnd.get_scc_precision()
    assert rc == 0 and settings == (0.0, 0, 0) and counters == dict(fp32_cycles=0, fp64_restart=0, scc_cycles=4)
    hnd.close()


def test_schedule_switches_once_and_reports(torch_cuda):
    torch = torch_cuda
    hnd = xtbeig.Handle(0)
    hnd.set_scc_precision(1e-3, 120, 0)
    h, s, sd, w, c = setup(torch, 64, 8)
    used, _ = run_sequence(torch, hnd, h, s, sd, w, c)
    norms = [1e-1 * (1e-6) ** (k / 19) for k in range(20)]
    expect = [xtbeig.FP32_REFINE if x >= 1e-3 else xtbeig.FP64 for x in norms]
    assert used == expect
    settings, counters, rc = hnd.get_scc_precision()
    assert rc == 0 and settings == (1e-3, 120, 0)
    assert counters == dict(fp32_cycles=sum(x >= 1e-3 for x in norms), fp64_restart=0, scc_cycles=20)
    # maxiter caps the FP32 cycles even while the change stays large; FP64 is kept to the end
    hnd.set_scc_precision(1.0, 2, 1)
    hnd.scc_begin()
    hd = torch.from_numpy(np.ascontiguousarray(h.T)).cuda()
    got = [hnd.scc_sygvd(hd, sd, w, c, 10.0) for _ in range(4)]
    assert got == [xtbeig.FP32_REFINE, xtbeig.FP32_REFINE, xtbeig.FP64, xtbeig.FP64]
    hnd.close()


@pytest.mark.parametrize("args,what", [((-1e-3, 160, 0), "negative fp32_rmsq"), ((float("nan"), 160, 0), "NaN"),
                                       ((float("inf"), 160, 0), "inf"), ((1e-3, 0, 0), "fp32_maxiter 0"),
                                       ((1e-3, 160, 2), "fp32_refine 2"), ((0.0, 160, 1), "refine without FP32")])
def test_invalid_settings_rejected_unchanged(torch_cuda, args, what):
    hnd = xtbeig.Handle(0)
    hnd.set_scc_precision(1e-2, 100, 0)
    with pytest.raises(ValueError) as e:
        hnd.set_scc_precision(*args)
    assert str(e.value)
    settings, _, _ = hnd.get_scc_precision()
    assert settings == (1e-2, 100, 0), what
    hnd.close()


def test_two_threads_bit_identical_to_serial(torch_cuda):
    torch = torch_cuda
    n = 96
    h, s, sd, _, _ = setup(torch, n, 9)

    def run(rmsq, out, key):
        hnd = xtbeig.Handle(0)
        hnd.set_scc_precision(rmsq, 160, 1 if rmsq > 0 else 0)
        w = torch.zeros(n, dtype=torch.float64, device="cuda")
        c = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        used, res = run_sequence(torch, hnd, h, s, sd, w, c, steps=6, seed=13)
        out[key] = (used, [(r[0], r[1]) for r in res])
        hnd.close()

    serial, threaded = {}, {}
    run(1e-3, serial, 32)
    run(0.0, serial, 64)
    ts = [threading.Thread(target=run, args=(1e-3, threaded, 32)), threading.Thread(target=run, args=(0.0, threaded, 64))]
    [t.start() for t in ts]
    [t.join() for t in ts]
    for key in (32, 64):
        assert threaded[key][0] == serial[key][0]
        for (w1, c1), (w2, c2) in zip(threaded[key][1], serial[key][1]):
            assert np.array_equal(w1, w2) and np.array_equal(c1, c2)
    assert xtbeig.FP32_REFINE in serial[32][0] and set(serial[64][0]) == {xtbeig.FP64}


@pytest.mark.parametrize("refine", [1, 0])
def test_sequence_error_against_the_oracle(torch_cuda, refine):
    """The configuration's sequence at n = 412, band 32: with refinement every step passes the FP64 gates; without
    it the FP32 steps stay within FP32 accuracy (eigenvalue error <= 1e-5 ||H||) and the FP64 steps pass."""
    torch = torch_cuda
    n = 412
    hnd = xtbeig.Handle(0)
    hnd.set_tuning(32, 2)
    hnd.set_scc_precision(1e-3, 160, refine)
    h0, s, sd, w, c = setup(torch, n, workloads.SEED_BASE + 2)
    used, res = run_sequence(torch, hnd, h0, s, sd, w, c, seed=workloads.SEED_BASE + 102)
    for u, (wg, cg, hk) in zip(used, res):
        g = workloads.gates(hk, s, wg, cg.T, sl.eigh(hk, s, eigvals_only=True))
        if refine or u == xtbeig.FP64:
            assert workloads.passes(g), (u, g)
        else:
            assert g["eig_err"] <= 1e-5, g
    hnd.close()
