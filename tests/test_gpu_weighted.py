"""Differentiable weighted PageRank on the GPU vs the FP64 oracle.

Forward: FP32 ranks within L1 <= 1e-6 of the FP64 weighted iteration at the
same sweep count.  Backward: dL/dw and dL/dt against the FP64 adjoint
(relative L2 error <= 1e-4) and against central finite differences of the
FP64 forward (h = 1e-3, 64 random edges), for L = sum_v wl_v r_v.
"""
import sys

import numpy as np
import pytest

import nwgr_amd
import oracle
from nwgr_amd import PageRankGraph
from tests.conftest import l1

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
MODES = {"leak": oracle.LEAK, "uniform": oracle.UNIFORM, "teleport": oracle.TELEPORT}
D = 0.85


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def case(scale, ef=16, seed=0):
    g = nwgr_amd.rmat(scale, ef)
    rng = np.random.default_rng(seed)
    w = (1.0 - rng.random(g.m)).astype(np.float32)           # (0, 1]
    t = rng.dirichlet(np.ones(g.n)).astype(np.float32)       # Dirichlet(1)
    wl = rng.standard_normal(g.n).astype(np.float32)         # loss weights
    return g, w, t, wl


def rel2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")


def test_unit_weights_equal_plain_solve():
    g = nwgr_amd.rmat(12, 16)
    deg = nwgr_amd.build_degrees(g)
    pg = PageRankGraph(g, deg)
    r_plain, res_p = pg.run(tol=1e-6, max_iters=100, dangling="uniform")
    w = torch.ones(g.m, dtype=torch.float32, device="cuda:0")
    out = torch.empty(g.n, dtype=torch.float32, device="cuda:0")
    res = pg.wpr_forward(w.data_ptr(), out.data_ptr(), tol=1e-6, max_iters=100, dangling="uniform")
    torch.cuda.synchronize()
    assert res.iters == res_p.iters
    np.testing.assert_array_equal(out.cpu().numpy(), r_plain)


@pytest.mark.parametrize("mode", ["leak", "uniform", "teleport"])
def test_weighted_forward_parity(mode):
    say(f"weighted forward s12 {mode}")
    g, w, t, _ = case(12, seed=1)
    pg = PageRankGraph(g)
    out = torch.empty(g.n, dtype=torch.float32, device="cuda:0")
    wd, td = dev(w), dev(t)
    res = pg.wpr_forward(wd.data_ptr(), out.data_ptr(), tol=1e-6, max_iters=200, dangling=mode,
                         teleport_ptr=td.data_ptr())
    torch.cuda.synchronize()
    r64, _ = oracle.wpr_forward(g.offsets, g.indices, w.astype(np.float64), t.astype(np.float64), D, 0.0,
                                res.iters, MODES[mode])
    _, k_ref = oracle.wpr_forward(g.offsets, g.indices, w.astype(np.float64), t.astype(np.float64), D, 1e-6,
                                  200, MODES[mode])
    assert l1(out.cpu().numpy(), r64) <= 1e-6
    assert abs(res.iters - k_ref) <= 1


@pytest.mark.parametrize("mode", ["leak", "uniform", "teleport"])
def test_autograd_matches_fp64_adjoint_and_fd(mode):
    say(f"autograd s14 {mode}")
    from nwgr_amd.autograd import weighted_pagerank
    g, w, t, wl = case(14, seed=2)
    pg = PageRankGraph(g)
    wd = dev(w).requires_grad_(True)
    td = dev(t).requires_grad_(True)
    r = weighted_pagerank(pg, wd, td, tol=1e-7, max_iters=300, dangling=mode, adj_tol=1e-7)
    loss = (r * dev(wl)).sum()
    loss.backward()
    gw, gt = wd.grad.cpu().numpy(), td.grad.cpu().numpy()
    w64, t64, wl64 = w.astype(np.float64), t.astype(np.float64), wl.astype(np.float64)
    r64, _ = oracle.wpr_forward(g.offsets, g.indices, w64, t64, D, 1e-15, 5000, MODES[mode], nthreads=8)
    _, dw64, dt64, _ = oracle.wpr_adjoint(g.offsets, g.indices, w64, t64, r64, wl64, D, 1e-13, 5000, MODES[mode],
                                          nthreads=8)
    ew, et = rel2(gw, dw64), rel2(gt, dt64)
    say(f"  rel err dw {ew:.2e} dt {et:.2e}")
    assert ew <= 1e-4 and et <= 1e-4
    # central finite differences of the FP64 forward, h = 1e-3, 64 random edges
    rng = np.random.default_rng(7)
    h = 1e-3
    ks = rng.choice(g.m, 64, replace=False)
    fd = np.empty(len(ks))
    for i, k in enumerate(ks):
        wp, wm = w64.copy(), w64.copy()
        wp[k] += h
        wm[k] -= h
        rp, _ = oracle.wpr_forward(g.offsets, g.indices, wp, t64, D, 1e-15, 5000, MODES[mode], nthreads=8, init=r64)
        rm, _ = oracle.wpr_forward(g.offsets, g.indices, wm, t64, D, 1e-15, 5000, MODES[mode], nthreads=8, init=r64)
        fd[i] = (wl64 @ rp - wl64 @ rm) / (2 * h)
    efd = rel2(gw[ks], fd)
    say(f"  rel err vs FD (64 edges) {efd:.2e}")
    assert efd <= 1e-4


@pytest.mark.parametrize("combine", [0, 1])
def test_weighted_gradients_low_locality_tiled(combine):
    """Forward and gradients through the tiled engine on a graph whose chunks
    hold more fragments than one LDS window (about one entry per row and
    2^13-wide tile), with the gathering and the accumulating second pass."""
    say(f"weighted low-locality tiled combine={combine}")
    from nwgr_amd.autograd import weighted_pagerank
    n = 1 << 16
    rng = np.random.default_rng(11)
    g = nwgr_amd.build_adjacency(rng.integers(0, n, 8 * n), rng.integers(0, n, 8 * n), n=n, dedup=True,
                                 drop_self_loops=True)
    w = (1.0 - rng.random(g.m)).astype(np.float32)
    t = rng.dirichlet(np.ones(g.n)).astype(np.float32)
    wl = rng.standard_normal(g.n).astype(np.float32)
    pg = PageRankGraph(g)
    pg.tuning = dict(engine=1, tile_bits=13, tile_fill_cost=0, tile_combine=combine)
    w64, t64 = w.astype(np.float64), t.astype(np.float64)
    # forward at the matched sweep count
    out = torch.empty(g.n, dtype=torch.float32, device="cuda:0")
    res = pg.wpr_forward(dev(w).data_ptr(), out.data_ptr(), tol=1e-6, max_iters=200, dangling="uniform",
                         teleport_ptr=dev(t).data_ptr())
    torch.cuda.synchronize()
    assert pg.last_engine().startswith("tiled")
    rk, _ = oracle.wpr_forward(g.offsets, g.indices, w64, t64, D, 0.0, res.iters, oracle.UNIFORM, nthreads=8)
    assert l1(out.cpu().numpy(), rk) <= 1e-6
    # gradients
    wd = dev(w).requires_grad_(True)
    td = dev(t).requires_grad_(True)
    r = weighted_pagerank(pg, wd, td, tol=1e-7, max_iters=300, dangling="uniform", adj_tol=1e-7)
    (r * dev(wl)).sum().backward()
    r64, _ = oracle.wpr_forward(g.offsets, g.indices, w64, t64, D, 1e-15, 5000, oracle.UNIFORM, nthreads=8)
    _, dw64, dt64, _ = oracle.wpr_adjoint(g.offsets, g.indices, w64, t64, r64, wl.astype(np.float64), D, 1e-13,
                                          5000, oracle.UNIFORM, nthreads=8)
    assert rel2(wd.grad.cpu().numpy(), dw64) <= 1e-4
    assert rel2(td.grad.cpu().numpy(), dt64) <= 1e-4


def test_adjoint_without_teleport_grad():
    from nwgr_amd.autograd import weighted_pagerank
    g, w, _, wl = case(10, seed=3)
    pg = PageRankGraph(g)
    wd = dev(w).requires_grad_(True)
    r = weighted_pagerank(pg, wd, None, tol=1e-7, max_iters=300, dangling="uniform")
    (r * dev(wl)).sum().backward()
    w64 = w.astype(np.float64)
    t64 = np.full(g.n, 1.0 / g.n)
    r64, _ = oracle.wpr_forward(g.offsets, g.indices, w64, t64, D, 1e-15, 5000, oracle.UNIFORM)
    _, dw64, _, _ = oracle.wpr_adjoint(g.offsets, g.indices, w64, t64, r64, wl.astype(np.float64), D, 1e-13, 5000,
                                       oracle.UNIFORM)
    assert rel2(wd.grad.cpu().numpy(), dw64) <= 1e-4


@pytest.mark.slow
@pytest.mark.timeout(600)
def test_config5_rmat_s20_gradients_and_cost():
    """R-MAT s20 ef16, w ~ U(0,1], t ~ Dirichlet(1), L = sum wl r: gradients vs
    the FP64 adjoint and vs central finite differences (64 edges), and
    forward+adjoint time vs forward time."""
    from nwgr_amd.autograd import weighted_pagerank
    say("config5 s20 generate")
    g, w, t, wl = case(20, seed=5)
    pg = PageRankGraph(g)
    wd = dev(w).requires_grad_(True)
    td = dev(t).requires_grad_(True)
    wld = dev(wl)
    for _ in range(2):  # warm (transpose built, graphs captured)
        wd.grad = td.grad = None
        r = weighted_pagerank(pg, wd, td, tol=1e-7, max_iters=300, adj_tol=1e-7)
        (r * wld).sum().backward()
    torch.cuda.synchronize()
    fwd = pg  # the handle records nothing per call; time with events
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    wd.grad = td.grad = None
    e0.record()
    r = weighted_pagerank(pg, wd, td, tol=1e-7, max_iters=300, adj_tol=1e-7)
    e1.record()
    (r * wld).sum().backward()
    e2.record()
    torch.cuda.synchronize()
    t_f, t_fb = e0.elapsed_time(e1), e0.elapsed_time(e2)
    say(f"config5 fwd {t_f:.2f} ms fwd+adj {t_fb:.2f} ms ratio {t_fb / t_f:.2f}")
    say("config5 oracle")
    w64, t64, wl64 = w.astype(np.float64), t.astype(np.float64), wl.astype(np.float64)
    r64, _ = oracle.wpr_forward(g.offsets, g.indices, w64, t64, D, 1e-14, 5000, oracle.LEAK, nthreads=16)
    _, dw64, dt64, _ = oracle.wpr_adjoint(g.offsets, g.indices, w64, t64, r64, wl64, D, 1e-12, 5000, oracle.LEAK,
                                          nthreads=16)
    ew, et = rel2(wd.grad.cpu().numpy(), dw64), rel2(td.grad.cpu().numpy(), dt64)
    say(f"config5 rel err dw {ew:.2e} dt {et:.2e}")
    assert ew <= 1e-4 and et <= 1e-4
    assert t_fb / t_f <= 2.5
    # central finite differences of the FP64 forward at this size (h = 1e-3,
    # 64 random edges), each perturbed solve warm-started from r64
    say("config5 finite differences (64 edges)")
    rng = np.random.default_rng(55)
    h = 1e-3
    ks = rng.choice(g.m, 64, replace=False)
    fd = np.empty(len(ks))
    for i, k in enumerate(ks):
        wp, wm = w64.copy(), w64.copy()
        wp[k] += h
        wm[k] -= h
        rp, _ = oracle.wpr_forward(g.offsets, g.indices, wp, t64, D, 1e-15, 5000, oracle.LEAK, nthreads=16, init=r64)
        rm, _ = oracle.wpr_forward(g.offsets, g.indices, wm, t64, D, 1e-15, 5000, oracle.LEAK, nthreads=16, init=r64)
        fd[i] = (wl64 @ rp - wl64 @ rm) / (2 * h)
    efd = rel2(wd.grad.cpu().numpy()[ks], fd)
    say(f"config5 rel err vs FD (64 edges) {efd:.2e}")
    assert efd <= 1e-4
