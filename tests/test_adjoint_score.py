"""CPU checks of the roughness estimator of the adjoint (float64, numpy).

The adjoint differentiates roughness through the score of the sampled GGX half
vector under the visible-normal distribution (bsdfs.h:27-72, Heitz 2018):
∂log p(h)/∂α = ∂log G1(v)/∂α + ∂log D(h)/∂α, with the closed forms used in
kernels_adjoint.hip.  Two facts pin them: the score has zero mean, and for a
smooth integrand the score-function derivative equals the derivative taken by
moving the samples with α under fixed random numbers."""
import numpy as np
import pytest


def vndf(alpha, v, u1, u2):
    """The sampling steps of bsdfs.h:27-72 without the rejection wrapper, in
    the shading frame (normal = +z)."""
    vh = np.array([alpha * v[0], alpha * v[1], v[2]])
    vh /= np.linalg.norm(vh)
    l2 = vh[0] ** 2 + vh[1] ** 2
    t1v = np.array([-vh[1], vh[0], 0.0]) / np.sqrt(l2) if l2 > 0 else np.array([1.0, 0.0, 0.0])
    t2v = np.cross(vh, t1v)
    s = 0.5 * (1.0 + vh[2])
    r = np.sqrt(u1)
    phi = 2.0 * np.pi * u2
    t1, t2 = r * np.cos(phi), r * np.sin(phi)
    t2 = (1.0 - s) * np.sqrt(np.maximum(0.0, 1.0 - t1 * t1)) + s * t2
    nz = np.sqrt(np.maximum(0.0, 1.0 - t1 * t1 - t2 * t2))
    nh = t1[:, None] * t1v + t2[:, None] * t2v + nz[:, None] * vh
    ne = np.stack([alpha * nh[:, 0], alpha * nh[:, 1], np.maximum(0.0, nh[:, 2])], 1)
    return ne / np.linalg.norm(ne, axis=1, keepdims=True)


def dlog_g1(c, a):
    s = np.sqrt(a * a + (1.0 - a * a) * c * c)
    return -a * (1.0 - c * c) / (s * (c + s))


def dlog_ndf(ndh, a):
    d = ndh * ndh * (a * a - 1.0) + 1.0
    return 2.0 / a - 4.0 * a * ndh * ndh / d


def g1(c, a):
    return 2.0 * c / (c + np.sqrt(a * a + (1.0 - a * a) * c * c))


def ndf(ndh, a):
    d = ndh * ndh * (a * a - 1.0) + 1.0
    return a * a / (np.pi * d * d)


def test_log_derivative_closed_forms():
    for a in (0.05, 0.3, 0.8):
        for c in (0.1, 0.5, 0.95):
            e = 1e-6
            num = (np.log(g1(c, a + e)) - np.log(g1(c, a - e))) / (2 * e)
            assert dlog_g1(c, a) == pytest.approx(num, rel=1e-6)
            num = (np.log(ndf(c, a + e)) - np.log(ndf(c, a - e))) / (2 * e)
            assert dlog_ndf(c, a) == pytest.approx(num, rel=1e-6)


@pytest.mark.parametrize("alpha", [0.05, 0.3, 0.7])
def test_score_matches_reparameterised_derivative(alpha):
    rng = np.random.default_rng(1)
    n = 1_000_000
    theta = 0.9
    v = np.array([np.sin(theta), 0.0, np.cos(theta)])
    u1, u2 = rng.random(n), rng.random(n)
    h = vndf(alpha, v, u1, u2)
    score = dlog_g1(v[2], alpha) + dlog_ndf(h[:, 2], alpha)
    se = score.std() / np.sqrt(n)
    assert abs(score.mean()) < 5 * se  # ∂/∂α ∫ p = 0

    def g(x):
        return np.exp(3.0 * x[:, 0]) + x[:, 2] ** 3

    e = 1e-5
    reparam = (g(vndf(alpha + e, v, u1, u2)).mean() - g(vndf(alpha - e, v, u1, u2)).mean()) / (2 * e)
    est = g(h) * score
    assert est.mean() == pytest.approx(reparam, abs=5 * est.std() / np.sqrt(n))
