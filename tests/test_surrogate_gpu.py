"""HIP kernels of the interval-wrench surrogate against the CPU oracle (oracle/surrogate.py):
the MFMA GEMM, loss, gradients and Adam steps of float32 training, the fused bfloat16
inference, and the knife-edge patch."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from blademaster_hip import Config, Solver  # noqa: E402
from blademaster_hip.mlp import Mlp  # noqa: E402
from oracle import surrogate as o  # noqa: E402

DEV = "cuda:0"
SHAPE = (96, 256, 6)


def _dev(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device=DEV)


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (17, 33, 5), (64, 64, 16), (100, 70, 129), (257, 256, 1000)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_gemm_matches_float64(M, N, K, ta, tb):
    rng = np.random.default_rng(M * 1000 + N + K)
    A = rng.normal(size=(K, M) if ta else (M, K)).astype(np.float32)
    B = rng.normal(size=(N, K) if tb else (K, N)).astype(np.float32)
    expected = (A.T if ta else A).astype(np.float64) @ (B.T if tb else B).astype(np.float64)
    net = Mlp(8, 16, 2, capacity=4, device=DEV)
    C = torch.full((M, N), np.nan, dtype=torch.float32, device=DEV)
    lda, ldb = A.shape[1], B.shape[1]
    code = net._lib.bm_mlp_gemm(*(net_ptr(t) for t in (_dev(A), _dev(B), C)), M, N, K, lda, ldb, N, ta, tb,
                                net._stream())
    assert code == 0
    got = C.cpu().numpy()
    # float32 products, float32 accumulation: |error| <= K eps sum |a b| (tolerance 1e-5 relative).
    scale = (np.abs(A.T if ta else A).astype(np.float64) @ np.abs(B.T if tb else B).astype(np.float64))
    assert np.all(np.abs(got - expected) <= 1e-5 * scale + 1e-30)


@pytest.mark.parametrize("M,N,K,work", [(257, 256, 4096, 16 * 257 * 256), (97, 256, 1000, 16 * 97 * 256),
                                         (257, 6, 4096, 3 * 257 * 6), (33, 17, 300, 0)])
def test_split_gemm_matches_float64(M, N, K, work):
    """Weight-gradient form (A transposed, K = batch), K split over workgroups."""
    rng = np.random.default_rng(M + N + K)
    A = rng.normal(size=(K, M)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    expected = A.T.astype(np.float64) @ B.astype(np.float64)
    net = Mlp(8, 16, 2, capacity=4, device=DEV)
    C = torch.full((M, N), np.nan, dtype=torch.float32, device=DEV)
    W = torch.zeros(max(work, 1), dtype=torch.float32, device=DEV)
    code = net._lib.bm_mlp_gemm_split(*(net_ptr(t) for t in (_dev(A), _dev(B), C)), M, N, K, M, N, 1, 0,
                                      net_ptr(W), work, net._stream())
    assert code == 0
    scale = np.abs(A.T).astype(np.float64) @ np.abs(B).astype(np.float64)
    assert np.all(np.abs(C.cpu().numpy() - expected) <= 1e-5 * scale + 1e-30)


def net_ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


def _net_and_data(batch=200, seed=0):
    rng = np.random.default_rng(seed)
    net = Mlp(*SHAPE, capacity=512, device=DEV, seed=seed)
    params = net.params.cpu().numpy().astype(np.float64)
    for w in o.split(params, *SHAPE):
        w[-1] = rng.normal(0, 0.1, w.shape[1])  # nonzero biases
    net.params.copy_(torch.tensor(params, dtype=torch.float32))
    x = rng.normal(size=(batch, SHAPE[0])).astype(np.float32)
    t = rng.normal(size=(batch, SHAPE[2])).astype(np.float32)
    return net, x, t


def test_forward_and_gradients_match_oracle():
    net, x, t = _net_and_data()
    params = net.params.cpu().numpy()
    loss = float(net.backward(_dev(x), _dev(t)).cpu()[0])
    h1, h2, out = o.forward(params, x, *SHAPE)
    value, grad = o.backward(params, x, t, *SHAPE)
    np.testing.assert_allclose(net.out[: len(x)].cpu().numpy(), out, rtol=1e-4, atol=1e-4 * np.abs(out).max())
    np.testing.assert_allclose(net.h1[: len(x)].cpu().numpy(), h1, rtol=1e-4, atol=1e-5 * np.abs(h1).max())
    assert loss == pytest.approx(value, rel=1e-5)
    got = net.grads.cpu().numpy()
    # Per layer, relative to the layer's largest gradient (float32 sums over the batch).
    o_ = 0
    for r, c in o.shapes(*SHAPE):
        sl = slice(o_, o_ + r * c)
        o_ += r * c
        np.testing.assert_allclose(got[sl], grad[sl], rtol=0, atol=2e-4 * np.abs(grad[sl]).max())
    assert int(net.step.cpu()[0]) == 0  # a gradient evaluation is not a training step


def test_adam_steps_match_oracle():
    net, x, t = _net_and_data(batch=128, seed=1)
    lr, decay = 1e-3, 0.05
    net.lr.copy_(torch.tensor([lr, decay]))
    params = net.params.cpu().numpy().astype(np.float64)
    m = np.zeros_like(params)
    v = np.zeros_like(params)
    xs, ts = _dev(x), _dev(t)
    for step in range(1, 4):
        _, grad = o.backward(params, x, t, *SHAPE)
        params, m, v = o.adam(params, grad, m, v, step, lr, weight_decay=decay)
        net.train_step(xs, ts)
    assert int(net.step.cpu()[0]) == 3
    got = net.params.cpu().numpy()
    # A gradient near zero can change sign between float32 and float64 (Adam then moves the
    # weight by up to 2 lr per step); elsewhere the weights agree to float32 rounding.
    steady = np.abs(o.backward(params, x, t, *SHAPE)[1]) > 1e-3 * np.abs(grad).max()
    assert steady.mean() > 0.5
    np.testing.assert_allclose(got[steady], params[steady], rtol=0, atol=2e-5)
    assert np.abs(got - params).max() <= 6 * lr + 1e-6


def test_bf16_inference_matches_oracle():
    net, x, _ = _net_and_data(batch=300, seed=2)
    rng = np.random.default_rng(3)
    shift = rng.normal(size=SHAPE[0]).astype(np.float32)
    scale = rng.uniform(0.5, 2.0, SHAPE[0]).astype(np.float32)
    y_shift = rng.normal(size=SHAPE[2]).astype(np.float32)
    y_scale = rng.uniform(0.5, 3.0, SHAPE[2]).astype(np.float32)
    for name, value in (("x_shift", shift), ("x_scale", scale), ("y_shift", y_shift), ("y_scale", y_scale)):
        getattr(net, name).copy_(torch.tensor(value))
    net._packed_current = False
    got = net.infer(_dev(x)).cpu().numpy()
    params = net.params.cpu().numpy()
    ref = o.infer_bf16(params, x, *SHAPE, x_shift=shift, x_scale=scale, y_shift=y_shift, y_scale=y_scale)
    # Same bf16 roundings; only the float32 accumulation order (and a hidden value rounded to
    # the neighbouring bf16 after it) differ.
    spread = np.abs(ref - y_shift).max(0)
    assert np.all(np.abs(got - ref) <= 2e-3 * spread + 1e-6)
    # And bf16 inference stays close to the float32 network.
    full = o.forward(params, (x - shift) * scale, *SHAPE)[2] * y_scale + y_shift
    assert np.abs(got - full).max() <= 0.05 * spread.max()


def test_fit_learns_a_smooth_map_and_graph_replays_equal_eager_steps():
    rng = np.random.default_rng(4)
    n = 8192
    X = rng.uniform(-1, 1, size=(n, SHAPE[0])).astype(np.float32)
    A = rng.normal(size=(SHAPE[0], SHAPE[2])) / np.sqrt(SHAPE[0])
    Y = (np.sin(2 * X @ A) + 0.5 * (X @ A) ** 2).astype(np.float32)
    Xd, Yd = _dev(X), _dev(Y)
    nets = []
    for graph in (True, False):
        net = Mlp(*SHAPE, capacity=1024, device=DEV, seed=5)
        net.standardize(Xd, Yd)
        history = net.fit(Xd, Yd, epochs=3, batch=512, lr=2e-3, seed=6, graph=graph)
        nets.append((net, history))
    (g_net, g_hist), (e_net, e_hist) = nets
    assert torch.equal(g_net.params, e_net.params)
    assert g_hist == e_hist
    net = g_net
    more = net.fit(Xd, Yd, epochs=12, batch=512, lr=2e-3, seed=7)
    assert more[-1] < 0.3 * g_hist[0]
    pred = net.infer(Xd).cpu().numpy()
    mse = np.mean((pred - Y) ** 2)
    # A nonlinear map of 96 inputs: the trained network explains over 80 % of its variance.
    assert mse < 0.2 * np.var(Y)


def test_blade_patch_matches_oracle():
    cfg = Config(size=(0.048, 0.024, 0.03), grid_spacing=0.003, dt=1e-4, external_tool=True, tool_substep="predict",
                 blade_half_size=(0.0005, 0.02, 0.01))
    s = Solver(cfg)
    rng = np.random.default_rng(8)
    x = s.reference + rng.normal(0, 3e-4, size=s.reference.shape)
    damage = rng.uniform(0, 1, s.n).astype(np.float32)
    s.x.assign(x.astype(np.float32))
    s.damage.assign(damage)
    ang = 0.2
    q = np.array([np.sin(ang / 2), 0.0, 0.0, np.cos(ang / 2)], dtype=np.float32)
    pos = np.array([0.0012, -0.001, 0.03 - 0.004 + 0.01], dtype=np.float32)
    s.tool_pos.assign(pos[None])
    s.tool_rot.assign(q[None])
    got = s.blade_patch().cpu().numpy()
    ref, margin = o.patch(x.astype(np.float32), damage, np.zeros(s.n, np.int64), pos[None].astype(np.float64),
                          q[None].astype(np.float64), cfg.grid_spacing, cfg.blade_half_size)
    assert ref[0, 0].sum() > 100
    # Particles within 1e-4 cell widths of a face may land on either side in float32.
    ambiguous = int((margin < 1e-4).sum())
    assert np.abs(got[:, 0] - ref[:, 0]).sum() <= 2 * ambiguous
    if ambiguous == 0:
        np.testing.assert_array_equal(got[:, 0], ref[:, 0])
        np.testing.assert_allclose(got[:, 1], ref[:, 1], rtol=1e-5, atol=1e-5)
