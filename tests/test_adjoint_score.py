"""CPU checks of the roughness estimator of the adjoint (float64, numpy).

The adjoint differentiates roughness through the score of the sampled GGX half
vector under the visible-normal distribution (bsdfs.h:27-72, Heitz 2018):
∂log p(h)/∂α = ∂log G1(v)/∂α + ∂log D(h)/∂α, with the closed forms used in
kernels_adjoint.hip.  Two facts pin them: the score has zero mean, and for a
smooth integrand the score-function derivative equals the derivative taken by
moving the samples with α under fixed random numbers."""
import math

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


def _vndf_numpy(alpha, mu, n, rng):
    """Heitz 2018's visible-normal sampler (bsdfs.h:27-72 steps) in float64:
    half vectors for the view (√(1−μ²), 0, μ), normal z."""
    v = np.array([math.sqrt(max(0.0, 1 - mu * mu)), 0.0, mu])
    vh = np.array([alpha * v[0], alpha * v[1], v[2]])
    vh /= np.linalg.norm(vh)
    l2 = vh[0] ** 2 + vh[1] ** 2
    T1 = np.array([-vh[1], vh[0], 0.0]) / math.sqrt(l2) if l2 > 0 else np.array([1.0, 0.0, 0.0])
    T2 = np.cross(vh, T1)
    s = 0.5 * (1 + vh[2])
    u1, u2 = rng.random(n), rng.random(n)
    r, phi = np.sqrt(u1), 2 * math.pi * u2
    t1, t2 = r * np.cos(phi), r * np.sin(phi)
    t2 = (1 - s) * np.sqrt(np.maximum(0.0, 1 - t1 * t1)) + s * t2
    nh = t1[:, None] * T1 + t2[:, None] * T2 + np.sqrt(np.maximum(0.0, 1 - t1 * t1 - t2 * t2))[:, None] * vh
    ne = np.stack([alpha * nh[:, 0], alpha * nh[:, 1], np.maximum(0.0, nh[:, 2])], 1)
    ne /= np.linalg.norm(ne, axis=1, keepdims=True)
    return v, ne


@pytest.mark.parametrize("alpha,mu", [(0.81, 0.3), (0.81, 0.8), (0.36, 0.15), (0.25, 0.6), (0.04, 0.05)])
def test_ggx_rejection_probability_and_derivative(alpha, mu):
    """phx_adjoint_ggx_rejection (the quadrature the roughness gradient uses
    for the rejection loop's acceptance) against a Monte-Carlo of the sampler
    itself: P = the fraction of draws whose mirror direction falls below the
    surface, ∂P/∂α = E[1{rejected}·∂log D_v(h)/∂α] (the region does not
    depend on α), within 5 standard errors."""
    import ctypes as C

    from phosphor_amd import lib

    out = np.zeros(2, np.float64)
    assert lib().phx_adjoint_ggx_rejection(C.c_float(alpha), C.c_float(mu), out.ctypes.data) == 0
    P, dP = out
    rng = np.random.default_rng(int(alpha * 1000 + mu * 100))
    n = 400_000
    v, h = _vndf_numpy(alpha, mu, n, rng)
    rej = (2 * (h @ v) * h[:, 2] - mu) < 0
    p_mc = rej.mean()
    assert abs(p_mc - P) < 5 * math.sqrt(max(P * (1 - P), 1e-12) / n) + 1e-4, (p_mc, P)
    a2 = alpha * alpha
    c = h[:, 2]
    d = c * c * (a2 - 1) + 1
    dlog_d = 2 / alpha - 4 * alpha * c * c / d
    root = math.sqrt(a2 + (1 - a2) * mu * mu)
    dlog_g1v = -alpha * (1 - mu * mu) / (root * (mu + root))
    x = rej * (dlog_d + dlog_g1v)
    assert abs(x.mean() - dP) < 5 * x.std() / math.sqrt(n) + 1e-4, (x.mean(), dP)
    # and the derivative is the slope of P in α
    lo, hi = np.zeros(2), np.zeros(2)
    e = 1e-3
    lib().phx_adjoint_ggx_rejection(C.c_float(alpha - e), C.c_float(mu), lo.ctypes.data)
    lib().phx_adjoint_ggx_rejection(C.c_float(alpha + e), C.c_float(mu), hi.ctypes.data)
    fd = (hi[0] - lo[0]) / (np.float32(alpha + e) - np.float32(alpha - e))
    assert abs(fd - dP) < 2e-3 * max(1.0, abs(dP)), (fd, dP)
