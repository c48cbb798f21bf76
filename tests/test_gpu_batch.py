"""Graph replay (xtbeig_dev_sygvd_graph), the batched uneven-n solve (xtbeig_dev_sygvd_batched) and the band
autotuner (xtbeig_autotune). Pins: every graph replay is bit-identical to the eager solve of the same input, with
eager work, other sizes and a workspace regrowth between replays; every batched result is bit-identical to the
eager solve of its problem on one stream and passes the FP64 gates against scipy; the tuned table is used and
round-trips."""
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
    hd, sd = dev(torch, h), dev(torch, s)  # referenced until the solve is done (it runs on the library's stream)
    hnd.sygvd(hd, sd, w, c, prec)
    assert hnd.synchronize() == [0, 0, 0, 0]
    return w.cpu().numpy(), c.cpu().numpy()


@pytest.mark.parametrize("prec", [xtbeig.FP64, xtbeig.FP32_REFINE])
def test_graph_replays_match_eager(torch_cuda, prec):
    torch = torch_cuda
    hnd = xtbeig.Handle(0)
    n = 131
    h, s = workloads.generalized_pair(n, 301)
    h2, _ = workloads.generalized_pair(n, 302)
    we, ce = solve_eager(torch, hnd, h, s, prec)
    we2, ce2 = solve_eager(torch, hnd, h2, s, prec)
    hd, sd, h2d = dev(torch, h), dev(torch, s), dev(torch, h2)
    w = torch.zeros(n, dtype=torch.float64, device="cuda")
    c = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    for rep in range(4):
        hnd.sygvd_graph(hd, sd, w, c, prec)
        assert hnd.synchronize() == [0, 0, 0, 0]
        assert np.array_equal(w.cpu().numpy(), we) and np.array_equal(c.cpu().numpy(), ce), rep
        # eager work of another size in between
        solve_eager(torch, hnd, *workloads.generalized_pair(64, 303 + rep), prec)
        hnd.sygvd_graph(h2d, sd, w, c, prec)
        assert hnd.synchronize() == [0, 0, 0, 0]
        assert np.array_equal(w.cpu().numpy(), we2) and np.array_equal(c.cpu().numpy(), ce2), rep
    # a larger solve regrows the workspace: the graph is captured again and still matches
    solve_eager(torch, hnd, *workloads.generalized_pair(300, 310), prec)
    hnd.sygvd_graph(hd, sd, w, c, prec)
    assert hnd.synchronize() == [0, 0, 0, 0]
    assert np.array_equal(w.cpu().numpy(), we) and np.array_equal(c.cpu().numpy(), ce)
    hnd.close()


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("prec", [xtbeig.FP64, xtbeig.FP32_REFINE])
def test_batched_matches_eager_and_oracle(torch_cuda, prec, use_graph):
    torch = torch_cuda
    rng = np.random.default_rng(1003)
    ns = [int(round(np.exp(x))) for x in rng.uniform(np.log(40), np.log(300), 12)]
    probs = [workloads.generalized_pair(n, 400 + q) for q, n in enumerate(ns)]
    hnd = xtbeig.Handle(0)
    hs = [dev(torch, h) for h, _ in probs]
    ss = [dev(torch, s) for _, s in probs]
    ws = [torch.zeros(n, dtype=torch.float64, device="cuda") for n in ns]
    cs = [torch.zeros((n, n), dtype=torch.float64, device="cuda") for n in ns]
    lat = hnd.sygvd_batched(hs, ss, ws, cs, prec, slots=4, use_graph=use_graph)
    assert hnd.synchronize() == [0, 0, 0, 0]
    assert len(lat) == len(ns) and all(x > 0 for x in lat)
    ref = xtbeig.Handle(0)
    for q, (h, s) in enumerate(probs):
        we, ce = solve_eager(torch, ref, h, s, prec)
        assert np.array_equal(ws[q].cpu().numpy(), we) and np.array_equal(cs[q].cpu().numpy(), ce), q
        g = workloads.gates(h, s, we, ce.T, sl.eigh(h, s, eigvals_only=True))
        assert workloads.passes(g), (q, g)
    ref.close()
    hnd.close()


def test_autotune_table(torch_cuda):
    torch = torch_cuda
    hnd = xtbeig.Handle(0)
    res = hnd.autotune(xtbeig.FP64, 100, 400)
    assert len(res) >= 2
    for start, band, ms in res:
        assert band in (8, 16, 32) and ms > 0
    starts, bands = hnd.band_table(xtbeig.FP64)
    for start, band, _ in res:
        assert bands[starts.index(start)] == band
    res32 = hnd.autotune(xtbeig.FP32_REFINE, 100, 400)
    assert all(b in (8, 16, 32) for _, b, _ in res32)
    # a solve after tuning passes the gates; the table round-trips
    h, s = workloads.generalized_pair(250, 501)
    w, c = solve_eager(torch, hnd, h, s, xtbeig.FP32_REFINE)
    assert workloads.passes(workloads.gates(h, s, w, c.T, sl.eigh(h, s, eigvals_only=True)))
    other = xtbeig.Handle(0)
    other.set_band_table(xtbeig.FP64, bands)
    assert other.band_table(xtbeig.FP64)[1] == bands
    with pytest.raises(RuntimeError):
        other.set_band_table(xtbeig.FP64, [65] * len(bands))
    other.close()
    hnd.close()
