"""Gradients of a render (phx_adjoint_*) against finite differences.

The backward replays every pass from its seed; the replay check asserts that
the replay kernels follow the forward's paths bit for bit (pixel emission and
direct light, deposits per photon).  Albedo and emission gradients are then
compared with central differences of the forward under common random numbers
(the same seeds), step 1e-2, two photon bounces (no photon roulette, whose
flips make difference quotients noisy) — tolerance 2 % relative per parameter
group, the figure the gradient is specified to.

Roughness gradients are not compared here: a roughness change alters how many
random numbers the GGX rejection loop consumes, which shifts the rest of the
random stream, so a difference quotient under common random numbers is
dominated by these jumps.  The roughness estimator (the half-vector score) is
checked in float64 by tests/test_adjoint_score.py."""
import numpy as np
import pytest

from phosphor_amd import PassConfig, scenes
from phosphor_amd.adjoint import PARAMS, Adjoint, l2_loss

pytestmark = pytest.mark.gpu


def _setup(res=128, photons=1 << 17, rounds=8, bounces=2):
    cfg = PassConfig(width=res, height=res, photons_per_round=photons, max_photons_per_batch=photons,
                     max_photon_bounces=bounces, grid_res=43, seed=3, initial_radius=0.02)
    scene = scenes.adjoint_scene(res, res, seed=1)
    adj = Adjoint(cfg, scene, max_rounds=rounds)
    th_true = adj.theta0()
    target = adj.forward(th_true, rounds)
    rng = np.random.default_rng(7)
    th = th_true.copy()
    th[:, :4] *= rng.uniform(0.8, 1.2, size=(th.shape[0], 4)).astype(np.float32)
    th[7, 4] = 1.1
    return adj, th, target, rounds


def test_adjoint_replay_follows_forward():
    adj, th, target, rounds = _setup(res=64, photons=1 << 14, rounds=3, bounces=8)
    img = adj.forward(th, rounds)
    _, g = l2_loss(img, target)
    adj.set_replay_check(True)
    grad = adj.backward(g)
    assert np.all(np.isfinite(grad))
    e, d, c, fc = adj.read_replay()
    hp = adj.replayed_hitpoints()
    assert np.array_equal(e.view(np.uint32), hp[:, 5].view(np.uint32)), "replayed emission differs"
    assert np.array_equal(d.view(np.uint32), hp[:, 4].view(np.uint32)), "replayed direct light differs"
    assert np.array_equal(c, fc) and fc.sum() > 0, "replayed deposits per photon differ"
    adj.close()


def test_adjoint_albedo_and_emission_match_finite_differences():
    adj, th, target, rounds = _setup()
    img = adj.forward(th, rounds)
    loss, g = l2_loss(img, target)
    assert loss > 0
    grad = adj.backward(g)
    h = 1e-2
    checked = [(m, k) for m in range(th.shape[0]) for k in range(3)] + [(7, 4)]
    fd = {}
    for m, k in checked:
        tp, tm = th.copy(), th.copy()
        tp[m, k] += h
        tm[m, k] -= h
        fd[(m, k)] = (l2_loss(adj.forward(tp, rounds), target)[0] - l2_loss(adj.forward(tm, rounds), target)[0]) / (2 * h)
    for k in (0, 1, 2, 4):
        keys = [mk for mk in checked if mk[1] == k]
        a = np.array([grad[m, kk] for m, kk in keys])
        f = np.array([fd[mk] for mk in keys])
        err = np.linalg.norm(a - f) / np.linalg.norm(f)
        print(f"{PARAMS[k]}: relative error {err:.4f}", flush=True)
        assert err <= 0.02, (PARAMS[k], a, f)
    adj.close()
