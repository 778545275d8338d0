"""CPU checks of the adjoint's host side: the torch.autograd wrapper routes
the image gradient to backward and returns its result, and the loss helper's
gradient is the derivative of the loss."""
import numpy as np
import pytest

from phosphor_amd.adjoint import l2_loss, render


class LinearStandIn:
    """image = A·θ: a render whose exact gradient is known."""

    def __init__(self, shape=(4, 5, 3), n_mat=2):
        rng = np.random.default_rng(0)
        self.shape = shape
        self.A = rng.standard_normal((int(np.prod(shape)), n_mat * 5)).astype(np.float32)
        self.n_mat = n_mat

    def forward(self, theta, rounds):
        return (self.A @ np.asarray(theta, np.float32).reshape(-1)).reshape(self.shape)

    def backward(self, dl):
        return (self.A.T @ np.asarray(dl, np.float32).reshape(-1)).reshape(self.n_mat, 5)


def test_l2_loss_gradient():
    rng = np.random.default_rng(2)
    img, tgt = rng.random((3, 4, 3)).astype(np.float32), rng.random((3, 4, 3)).astype(np.float32)
    loss, g = l2_loss(img, tgt)
    e = 1e-3
    d = np.zeros_like(img)
    d[1, 2, 0] = e
    num = (l2_loss(img + d, tgt)[0] - l2_loss(img - d, tgt)[0]) / (2 * e)
    assert g[1, 2, 0] == pytest.approx(num, rel=1e-3)


def test_autograd_function_routes_gradients():
    torch = pytest.importorskip("torch")
    fake = LinearStandIn()
    fn = render(fake, rounds=1)
    theta = torch.tensor(np.random.default_rng(1).random((2, 5)).astype(np.float32), requires_grad=True)
    target = torch.zeros(fake.shape)
    img = fn(theta)
    loss = ((img - target) ** 2).mean()
    loss.backward()
    dl = (2.0 * img.detach().numpy() / img.numel())
    expect = fake.backward(dl)
    assert np.allclose(theta.grad.numpy(), expect, rtol=1e-5, atol=1e-6)
