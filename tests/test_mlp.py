"""Warm-start MLP: the numpy oracle pinned by finite differences and by torch's
bfloat16 rounding (CPU), then the HIP GEMM / GELU / MSE / Adam kernels against
the oracle on the GPU."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo

SMALL = (7, 6, 5, 3)


def _fd_check(params, X, T):
    loss, grads = mo.loss_and_grads(params, X, T, bf16=False)
    rng = np.random.default_rng(3)
    for pi, (p, g) in enumerate(zip(params, grads)):
        for _ in range(4):
            idx = tuple(rng.integers(0, s) for s in p.shape)
            h = 1e-6
            pp = [q.copy() for q in params]
            pm = [q.copy() for q in params]
            pp[pi][idx] += h
            pm[pi][idx] -= h
            lp = np.mean((mo.forward(pp, X, bf16=False)[0] - T) ** 2)
            lm = np.mean((mo.forward(pm, X, bf16=False)[0] - T) ** 2)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)), (pi, idx, fd, g[idx])


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    params = [np.asarray(p, np.float64) for p in mo.init_params(SMALL, seed=2)]
    params[1] += rng.standard_normal(params[1].shape) * 0.1
    params[3] += rng.standard_normal(params[3].shape) * 0.1
    X = rng.standard_normal((9, SMALL[0]))
    T = rng.standard_normal((9, SMALL[-1]))
    _fd_check(params, X, T)


def test_oracle_gelu_derivative_and_values():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - mo.gelu_grad(z))) < 1e-8
    # known limits: gelu(0) = 0, gelu(x) -> x for large x, -> 0 for very negative x
    assert mo.gelu(0.0) == 0.0
    assert abs(mo.gelu(10.0) - 10.0) < 1e-12 and abs(mo.gelu(-10.0)) < 1e-12


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(4096).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 4096),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38, -1e-40], np.float32)])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.bf16_round(x), ref)


# ----------------------------------------------------------------- GPU
def _cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _bf16_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda().to(torch.bfloat16)


@pytest.mark.gpu
@pytest.mark.parametrize("tA,tB,M,N,K", [(0, 0, 64, 64, 16), (0, 0, 100, 70, 45), (1, 0, 33, 130, 77),
                                         (0, 1, 129, 31, 200), (1, 1, 17, 65, 19), (0, 0, 1024, 512, 769)])
def test_gemm_bf16_matches_numpy(tA, tB, M, N, K):
    torch = _cuda()
    from cochlearmodel_amd._lib import check, lib, ptr
    print(f"gemm tA={tA} tB={tB} M={M} N={N} K={K}", flush=True)
    rng = np.random.default_rng(M * 7 + N)
    A = rng.standard_normal((K, M) if tA else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tB else (K, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    dA, dB = _bf16_dev(torch, A), _bf16_dev(torch, B)
    dbias = torch.from_numpy(bias).cuda()
    dC = torch.from_numpy(C0).cuda()
    L = lib()
    check(L.cm_gemm_bf16(tA, tB, M, N, K, ptr(dA), A.shape[1], ptr(dB), B.shape[1], ptr(dC), N, ptr(dbias), 0.5))
    torch.cuda.synchronize()
    Ar, Br = mo.bf16_round(A), mo.bf16_round(B)
    ref = (Ar.T if tA else Ar) @ (Br.T if tB else Br) + bias + 0.5 * C0
    got = dC.cpu().numpy().astype(np.float64)
    scale = np.sqrt(K) * 4.0
    assert np.max(np.abs(got - ref)) <= 1e-5 * scale, np.max(np.abs(got - ref))


@pytest.mark.gpu
def test_mlp_forward_backward_match_oracle():
    torch = _cuda()
    from cochlearmodel_amd.mlp import SIZES, WarmStartMLP
    rng = np.random.default_rng(5)
    B = 96
    params = mo.init_params(SIZES, seed=4)
    params[1] = (rng.standard_normal(SIZES[1]) * 0.1).astype(np.float32)
    params[3] = (rng.standard_normal(SIZES[2]) * 0.1).astype(np.float32)
    X = rng.standard_normal((B, SIZES[0])).astype(np.float32)
    T = rng.standard_normal((B, SIZES[-1])).astype(np.float32)
    net = WarmStartMLP(device="cuda", params=params)
    Y = net.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    loss = net.backward(torch.from_numpy(T).cuda(), B)
    torch.cuda.synchronize()
    Yr, _ = mo.forward(params, X, bf16=True)
    lr, gr = mo.loss_and_grads(params, X, T, bf16=True)
    assert np.max(np.abs(Y - Yr)) <= 2e-3 * np.max(np.abs(Yr)), np.max(np.abs(Y - Yr))
    assert abs(loss - lr) <= 1e-4 * lr
    got = []
    for w, b in zip(net.gW, net.gb):
        got += [w.cpu().numpy(), b.cpu().numpy()]
    for i, (g, r) in enumerate(zip(got, gr)):
        rel = np.linalg.norm(g - r) / np.linalg.norm(r)
        print(f"grad {i}: rel {rel:.2e}", flush=True)
        assert rel <= 1e-2, (i, rel)


def test_oracle_adam_first_step_closed_form():
    # after one step the bias-corrected moments are g and g^2: w -= lr g / (|g| + eps)
    rng = np.random.default_rng(2)
    w, g = rng.standard_normal(50), rng.standard_normal(50)
    w1, _, _ = mo.adam_step(w, g, 0 * w, 0 * w, 1, 1e-3, eps=1e-8)
    assert np.allclose(w1, w - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)


@pytest.mark.gpu
def test_adam_kernel_matches_oracle():
    torch = _cuda()
    from cochlearmodel_amd._lib import check, lib, ptr
    rng = np.random.default_rng(4)
    n = 3000
    w = rng.standard_normal(n).astype(np.float32)
    m1 = np.zeros(n, np.float32)
    m2 = np.zeros(n, np.float32)
    dw, dm1, dm2 = (torch.from_numpy(a.copy()).cuda() for a in (w, m1, m2))
    dwb = torch.empty(n, dtype=torch.bfloat16, device="cuda")
    rw, rm1, rm2 = w.astype(np.float64), m1.astype(np.float64), m2.astype(np.float64)
    for step in (1, 2, 3):
        g = rng.standard_normal(n).astype(np.float32)
        dg = torch.from_numpy(g).cuda()
        check(lib().cm_adam(ptr(dw), ptr(dwb), ptr(dg), ptr(dm1), ptr(dm2), n, 1e-2, 0.9, 0.999, 1e-8, step, 1e-3))
        rw, rm1, rm2 = mo.adam_step(rw, g.astype(np.float64), rm1, rm2, step, 1e-2, wd=1e-3)
    torch.cuda.synchronize()
    got = dw.cpu().numpy()
    assert np.max(np.abs(got - rw)) <= 1e-6, np.max(np.abs(got - rw))
    assert np.array_equal(dwb.to(torch.float32).cpu().numpy().astype(np.float64), mo.bf16_round(got))


@pytest.mark.gpu
def test_mlp_adam_training_reduces_loss():
    torch = _cuda()
    from cochlearmodel_amd.mlp import SIZES, WarmStartMLP
    rng = np.random.default_rng(9)
    n = 4096
    X = rng.standard_normal((n, SIZES[0])).astype(np.float32)
    Wt = (rng.standard_normal((SIZES[0], SIZES[-1])) / np.sqrt(SIZES[0])).astype(np.float32)
    T = np.tanh(X @ Wt).astype(np.float32)
    net = WarmStartMLP(seed=1, device="cuda")
    hist = net.train(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda(), epochs=4, batch=256, lr=1e-3)
    print("train mse", hist, flush=True)
    assert all(b < a for a, b in zip(hist, hist[1:])) and hist[-1] < 0.8 * hist[0], hist
