"""The selector network's HIP forward / backward / Adam kernels against the float64 CPU oracle
(oracle/mlp_oracle.py).  FP32 kernels: tolerance rtol 1e-4, atol 1e-5 on forward values and
gradients (float32 accumulation over a batch of 512)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_mlp_forward_backward_matches_oracle(hx_gpu):
    from oracle import mlp_oracle as MO
    from hacs_amd import selector as S
    rng = np.random.default_rng(0)
    B, F, H, O = 512, 9, 64, 6
    X = rng.normal(size=(B, F)).astype(np.float32)
    T = rng.normal(size=(B, O)).astype(np.float32)
    mask = (rng.random((B, O)) < 0.7).astype(np.float32)
    net = S.MLP(F, H, O, seed=3)
    W1, b1, W2, b2 = [t.detach().cpu().numpy().astype(np.float64) for t in net.params()]
    hid, y = net.forward(torch.as_tensor(X, device="cuda"))
    h_ref, y_ref = MO.forward(X.astype(np.float64), W1, b1, W2, b2)
    np.testing.assert_allclose(y.cpu().numpy(), y_ref, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(hid.cpu().numpy(), h_ref, rtol=1e-4, atol=1e-5)
    loss, grads = net.backward(torch.as_tensor(X, device="cuda"), torch.as_tensor(T, device="cuda"),
                               torch.as_tensor(mask, device="cuda"))
    l_ref, *g_ref = MO.backward(X.astype(np.float64), W1, b1, W2, b2, T.astype(np.float64), mask.astype(np.float64))
    assert abs(loss - l_ref) <= 1e-4 * abs(l_ref) + 1e-6
    for g, gr in zip(grads, g_ref):
        np.testing.assert_allclose(g.cpu().numpy(), gr, rtol=1e-4, atol=1e-5)


def test_mlp_training_reduces_loss(hx_gpu):
    """Adam on the GPU fits a smooth target: the loss drops by more than 10x."""
    from hacs_amd import selector as S
    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, size=(1024, 4)).astype(np.float32)
    T = np.stack([X[:, 0] * X[:, 1], np.sin(2 * X[:, 2]), X[:, 3] ** 2], 1).astype(np.float32)
    net = S.MLP(4, 64, 3, seed=5)
    Xd, Td = torch.as_tensor(X, device="cuda"), torch.as_tensor(T, device="cuda")
    l0 = net.backward(Xd, Td)[0]
    for _ in range(400):
        net.train_step(Xd, Td, lr=1e-2)
    l1 = net.backward(Xd, Td)[0]
    assert l1 < l0 / 10, (l0, l1)
