"""Learned nest surrogate (csrc/surrogate.hip, cliffs-src_amd/surrogate.py).

CPU: the checkers themselves -- BF16 rounding against torch's, the written-
out backpropagation of oracle/surrogate_ref.py against torch autograd, the
quadtree layout, the gate.  GPU: the MFMA GEMM against a CPU product of the
same BF16 operands, the network's loss and gradients against both
restatements, training, and solver-generated data.
"""
import numpy as np
import pytest

import surrogate as S
import surrogate_ref as R


def _batch(rng, B):
    t = np.linspace(0, 1, S.T)[None, :, None]
    ph = rng.uniform(0, 2 * np.pi, (B, 1, S.G))
    x = (np.sin(2 * np.pi * 3 * t + ph) * rng.uniform(0.2, 1.0, (B, 1, S.G))).astype(np.float32)
    y = rng.normal(0, 0.3, (B, S.NOUT)).astype(np.float32)
    return x, y


def _params(seed=0):
    rng = np.random.default_rng(seed)
    ps = []
    for _, i, n in R.LAYERS:
        ps += [rng.uniform(-1, 1, i * n) * np.sqrt(6.0 / i), rng.normal(0, 0.05, n)]
    return np.concatenate(ps).astype(np.float32)


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-20, 100), [0.0, -0.0, 1.0, -3.5]])
    x = x.astype(np.float32)
    want = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(R.bf16(x).view(np.uint32), want.view(np.uint32))


def _per_param(g_a, g_b):
    """(relative L2 error, cosine) of two gradient vectors, per weight and bias."""
    out, o = [], 0
    for _, i, n in R.LAYERS:
        for k in (i * n, n):
            a, b = g_a[o:o + k].astype(np.float64), g_b[o:o + k].astype(np.float64)
            o += k
            nb = np.linalg.norm(b) + 1e-12
            out.append((np.linalg.norm(a - b) / nb, a @ b / (np.linalg.norm(a) * nb + 1e-30)))
    return out


def test_reference_backprop_matches_autograd():
    rng = np.random.default_rng(1)
    x, y = _batch(rng, 2)
    p = _params(1)
    l0, g0 = R.loss_grad_bf16(p, x, y, rounding=False)
    l2, g2 = R.loss_grad_autograd(p, x, y)
    assert abs(l0 - l2) <= 1e-5 * l2
    assert all(err <= 1e-4 for err, _ in _per_param(g0, g2))      # the written-out backprop
    l1, g1 = R.loss_grad_bf16(p, x, y)                               # with BF16 operands
    assert abs(l1 - l2) <= 0.02 * l2
    assert all(err <= 0.1 and cos >= 0.995 for err, cos in _per_param(g1, g2))


def test_quadtree_layout():
    q = S.quadtree_order().reshape(S.MAP, S.MAP)
    assert sorted(q.ravel()) == list(range(S.NOUT))
    # each 2x2 block is 4 consecutive outputs, each 8x8 block 64 consecutive
    for by in range(0, 64, 2):
        for bx in range(0, 64, 2):
            blk = q[by:by + 2, bx:bx + 2].ravel()
            assert blk.min() % 4 == 0 and blk.max() - blk.min() == 3
    for by in range(0, 64, 8):
        for bx in range(0, 64, 8):
            blk = q[by:by + 8, bx:bx + 8].ravel()
            assert blk.min() % 64 == 0 and blk.max() - blk.min() == 63
    maps = np.random.default_rng(2).normal(size=(3, 64, 64)).astype(np.float32)
    assert np.array_equal(S.from_quadtree(S.to_quadtree(maps)), maps)


def test_gate():
    a = np.ones((3, 64, 64))
    b = a.copy()
    b[1] *= 1.5
    b[2] = 0
    assert list(S.gate(a, b, tol=0.1)) == [True, False, False]
    # more networks: the spread of their run-ups against the mean
    c = a.copy()
    c[0] *= 1.05
    assert list(S.gate(a, a, c, tol=0.1)) == [True, True, True]
    assert list(S.gate(a, a, c, tol=0.01)) == [False, True, True]


@pytest.mark.gpu
def test_gemm_bf16_matches_reference(gpu_lib):
    rng = np.random.default_rng(3)
    for M, N, K in [(70, 33, 45), (128, 512, 64), (1, 4096, 512), (300, 4, 16), (64, 64, 2048)]:
        A = rng.normal(size=(M, K)).astype(np.float32)
        B = rng.normal(size=(K, N)).astype(np.float32)
        got = S.gemm_bf16(A, B)
        want = R.bf16(A).astype(np.float64) @ R.bf16(B).astype(np.float64)
        err = np.abs(got - want).max()
        assert err <= 2e-6 * K * max(1.0, np.abs(want).max()), (M, N, K, err)


@pytest.mark.gpu
def test_loss_and_gradients_match_references(gpu_lib):
    rng = np.random.default_rng(4)
    x, y = _batch(rng, 3)
    p = _params(2)
    net = S.Surrogate(batch_max=4, seed=0)
    net.set_params(p)
    loss, g = net.loss_grad(x, y)
    l_b, g_b = R.loss_grad_bf16(p, x, y)
    l_a, g_a = R.loss_grad_autograd(p, x, y)
    assert abs(loss - l_b) <= 1e-3 * l_b
    for l, ((err_b, _), (err_a, cos_a)) in enumerate(zip(_per_param(g, g_b), _per_param(g, g_a))):
        assert err_b <= 1e-2, (l, err_b)                 # same BF16 rounding points
        assert err_a <= 0.1 and cos_a >= 0.995, (l, err_a, cos_a)   # FP32 autograd
    out = net.forward(x)
    assert out.shape == (3, S.NOUT)
    net.close()


@pytest.mark.gpu
def test_training_reduces_loss(gpu_lib):
    rng = np.random.default_rng(5)
    x, _ = _batch(rng, 96)
    # a learnable target: a fixed smooth map per input, scaled by the series' energy
    base = np.sin(np.linspace(0, 3 * np.pi, S.NOUT)).astype(np.float32)
    y = (np.sqrt((x ** 2).mean(axis=(1, 2)))[:, None] * base[None, :]).astype(np.float32)
    net = S.Surrogate(batch_max=32, seed=1)
    losses = net.train(x, y, batch=32, steps=300, lr=1e-3, seed=3)
    assert np.isfinite(losses).all()
    assert losses[-20:].mean() < 0.1 * losses[:5].mean(), (losses[:5], losses[-5:])
    net.close()


@pytest.mark.gpu
def test_solver_data(gpu_lib):
    x, yq, maps, runups = S.dataset([0, 1])
    assert x.shape == (2, S.T, S.G) and yq.shape == (2, S.NOUT) and maps.shape == (2, 64, 64)
    assert np.isfinite(x).all() and (maps >= 0).all() and (runups >= 0).all()
    assert np.abs(x).max() > 0.0
    assert np.array_equal(S.from_quadtree(yq) / S.Y_SCALE, maps)
    # the ensemble's gauge series are the parent's own (cliffs_run_gauges)
    st = S.site()
    case = S.scenario(1)
    s1, s2, zeta = case.metrics()
    g = gpu_lib.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    g.upload(case.cel, case.xvel, case.yvel)
    alone = g.run_gauges(S.EVERY, S.T, st.gI, st.gJ)
    g.close()
    assert np.array_equal(alone * np.float32(S.X_SCALE), x[1])


@pytest.mark.gpu
def test_trained_network_gain(gpu_lib):
    """What the trained network gains over the method without it, on
    scenarios it has not seen: its maps are closer to the full nested solve
    than the training mean map (the best answer without a network), and
    ensemble parents plus inference cost a small fraction of the nested
    solve."""
    import time
    tr = S.dataset(range(64))
    cases = [S.scenario(s) for s in range(300_000, 300_032)]
    t0 = time.perf_counter()
    x = S.parent_gauges(cases) * S.X_SCALE
    t_par = (time.perf_counter() - t0) / len(cases)
    t0 = time.perf_counter()
    truth = [S.nested_solve(c) for c in cases]
    t_nest = (time.perf_counter() - t0) / len(cases)
    maps = np.array([m for m, _ in truth])
    net = S.Surrogate(batch_max=64, seed=1)
    losses = net.train(tr[0], tr[1], batch=32, steps=600, lr=5e-4, seed=2)
    assert np.isfinite(losses).all()
    t0 = time.perf_counter()
    pred = S.predict_maps(net, x)
    t_inf = (time.perf_counter() - t0) / len(cases)
    net.close()
    r_net = float(np.sqrt(((pred - maps) ** 2).mean()))
    r_mean = float(np.sqrt(((tr[2].mean(axis=0)[None] - maps) ** 2).mean()))
    print(f"map RMSE {r_net:.3f} m vs mean map {r_mean:.3f} m; per scenario: parents {t_par * 1e3:.1f} ms, "
          f"inference {t_inf * 1e3:.2f} ms, nested solve {t_nest * 1e3:.0f} ms")
    assert r_net < 0.95 * r_mean, (r_net, r_mean)
    assert t_par + t_inf < 0.2 * t_nest, (t_par, t_inf, t_nest)


@pytest.mark.gpu
def test_gate_falls_back_to_nested_solve(gpu_lib):
    """surrogate.answer: a scenario the gate rejects is answered by the full
    nested solve, bit for bit; an accepted one by the networks' mean."""
    cases = [S.scenario(s) for s in (400_000, 400_001)]
    nets = [S.Surrogate(batch_max=8, seed=k) for k in (1, 2)]
    maps, ok = S.answer(cases, nets, tol=-1.0)         # a negative tolerance accepts nothing
    assert not ok.any()
    for c, m in zip(cases, maps):
        assert np.array_equal(m, S.nested_solve(c)[0])
    maps, ok = S.answer(cases, nets, tol=np.inf)
    assert ok.all()
    x = S.parent_gauges(cases) * S.X_SCALE
    want = 0.5 * (S.predict_maps(nets[0], x) + S.predict_maps(nets[1], x))
    assert np.array_equal(maps, want)
    for n in nets:
        n.close()
