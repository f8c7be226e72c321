"""GPU parity of the cost-model network kernels (opensycl_amd/csrc/mlp.hip)
against the FP64 oracle (oracle/mlp.py): forward, loss, gradient, Adam step,
determinism, and a training run that must fit a smooth target."""
import ctypes

import numpy as np
import pytest

import opensycl_amd as acpp
from opensycl_amd import _abi
import oracle.mlp as M

pytestmark = pytest.mark.gpu


def _setup(torch, seed=0, B=1000):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((B, M.IN)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2]).astype(np.float32)
    p = M.init_params(seed)
    p[M.OFF_B1:M.OFF_W2] = (rng.standard_normal(M.H) * 0.1).astype(np.float32)
    p[M.OFF_B3] = 0.2
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return p, x, t, d(p), d(x), d(t)


def test_forward_matches_oracle(gpu, q):
    torch, _ = gpu
    p, x, t, pd, xd, td = _setup(torch, B=1537)
    yd = torch.empty(x.shape[0], dtype=torch.float32, device="cuda")
    _abi.check(_abi.load().acpp_mlp_forward(q.handle, pd.data_ptr(), xd.data_ptr(), x.shape[0], yd.data_ptr()))
    q.wait()
    y_ref, _, _ = M.forward(p, x)
    np.testing.assert_allclose(yd.cpu().numpy(), y_ref, rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("B", [1, 255, 256, 1000, 4099])
def test_gradient_and_loss_match_oracle(gpu, q, scratch, B):
    torch, _ = gpu
    p, x, t, pd, xd, td = _setup(torch, seed=B, B=B)
    grad = torch.empty(M.P, dtype=torch.float32, device="cuda")
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    _abi.check(_abi.load().acpp_mlp_gradient(scratch.handle, q.handle, pd.data_ptr(), xd.data_ptr(),
                                             td.data_ptr(), B, grad.data_ptr(), loss.data_ptr()))
    q.wait()
    L, g = M.loss_and_grad(p, x, t)
    assert abs(loss.item() - L) <= 1e-4 * abs(L) + 1e-6
    got = grad.cpu().numpy().astype(np.float64)
    # fp32 accumulation over at most 256-sample chunks: error bounded
    # relative to the largest gradient entry
    scale = np.abs(g).max()
    assert np.max(np.abs(got - g)) <= 2e-4 * scale + 1e-7, np.max(np.abs(got - g)) / scale


def test_gradient_is_deterministic(gpu, q, scratch):
    torch, _ = gpu
    p, x, t, pd, xd, td = _setup(torch, seed=5, B=3000)
    g1 = torch.empty(M.P, dtype=torch.float32, device="cuda")
    g2 = torch.empty_like(g1)
    lib = _abi.load()
    for g in (g1, g2):
        _abi.check(lib.acpp_mlp_gradient(scratch.handle, q.handle, pd.data_ptr(), xd.data_ptr(), td.data_ptr(),
                                         x.shape[0], g.data_ptr(), None))
        scratch.release()
    q.wait()
    assert torch.equal(g1, g2)


def test_adam_step_matches_oracle(gpu, q, scratch):
    torch, _ = gpu
    p, x, t, pd, xd, td = _setup(torch, seed=7, B=800)
    rng = np.random.default_rng(8)
    m0 = (rng.standard_normal(M.P) * 1e-3).astype(np.float32)
    v0 = (np.abs(rng.standard_normal(M.P)) * 1e-5).astype(np.float32)
    md, vd = torch.from_numpy(m0).cuda(), torch.from_numpy(v0).cuda()
    step, lr, wd = 3, 1e-3, 1e-2
    _abi.check(_abi.load().acpp_mlp_adam_step(scratch.handle, q.handle, pd.data_ptr(), md.data_ptr(), vd.data_ptr(),
                                              step, xd.data_ptr(), td.data_ptr(), x.shape[0], ctypes.c_float(lr),
                                              ctypes.c_float(wd), None))
    q.wait()
    _, g = M.loss_and_grad(p, x, t)
    p_ref, m_ref, v_ref = M.adam(p, m0, v0, step, g, lr, wd)
    np.testing.assert_allclose(md.cpu().numpy(), m_ref, rtol=1e-3, atol=1e-7)
    np.testing.assert_allclose(vd.cpu().numpy(), v_ref, rtol=1e-3, atol=1e-10)
    # each parameter moves by at most ~lr; compare the moves
    np.testing.assert_allclose(p - pd.cpu().numpy(), p - p_ref, rtol=2e-2, atol=2e-5 * lr / 1e-3)


def test_training_fits_smooth_target(gpu):
    torch, _ = gpu
    rng = np.random.default_rng(11)
    X = rng.standard_normal((2000, M.IN))
    y = np.exp(0.5 * np.sin(X[:, 0]) + 0.3 * X[:, 1] * X[:, 2] + 2.0)  # positive "times"
    m = acpp.CostModel(seed=1)
    hist = m.fit(X, y, steps=1500, lr=3e-3, log_every=100)
    assert hist[-1] < 0.1 * hist[0], hist
    pred = m.predict_log(X)
    resid = pred - np.log(y)
    assert np.sqrt(np.mean(resid ** 2)) < 0.1 * np.std(np.log(y)) + 0.05


def test_mlp_rejects_bad_arguments(gpu, q, scratch):
    torch, _ = gpu
    lib = _abi.load()
    assert lib.acpp_mlp_forward(q.handle, None, None, 1, None) == _abi.ERROR_INVALID_ARGUMENT
    p = torch.zeros(M.P, device="cuda")
    x = torch.zeros(16, device="cuda")
    assert lib.acpp_mlp_forward(q.handle, p.data_ptr(), x.data_ptr(), 0, x.data_ptr()) == \
        _abi.ERROR_INVALID_ARGUMENT
    # step 0 is rejected (bias correction counts from 1)
    assert lib.acpp_mlp_adam_step(scratch.handle, q.handle, p.data_ptr(), p.data_ptr(), p.data_ptr(), 0,
                                  x.data_ptr(), x.data_ptr(), 1, ctypes.c_float(1e-3), ctypes.c_float(0.0),
                                  None) == _abi.ERROR_INVALID_ARGUMENT
