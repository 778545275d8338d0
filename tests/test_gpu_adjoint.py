"""Gradients of a render (phx_adjoint_*) against finite differences.

The backward replays every pass from its seed; the replay check asserts that
the replay kernels follow the forward's paths bit for bit (pixel emission and
direct light, deposits per photon).  Albedo and emission gradients are then
compared with central differences of the forward under common random numbers
(the same seeds), step 1e-2, two photon bounces (no photon roulette, whose
flips make difference quotients noisy) — tolerance 2 % relative per parameter
group, the figure the gradient is specified to.

Roughness changes how many random numbers the GGX rejection loop consumes,
which shifts the rest of a path's random stream, so a single difference
quotient under common random numbers is dominated by these jumps.  Roughness
gradients are therefore compared with finite differences of the expectation:
a linear loss (whose gradient is the expectation of the estimator's), both
sides averaged over many independent seeds, within their statistical errors.
The score terms themselves (VNDF density, rejection loop) are checked in
float64 by tests/test_adjoint_score.py.

At the specified size (512², 2^20 photons, 32 passes, 16 bounces, roulette
active) albedo and emission gradients are compared with central differences
at steps 1e-2 and 3e-3 (within 2 %); at step 1e-3 the difference quotient's
deviation grows as 1/h — the photon powers are RGBE-quantised and roulette
decisions flip with the albedo, both piecewise constant in θ — so that step
measures the quotient's noise, not the gradient (printed, not asserted)."""
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


def test_adjoint_roughness_matches_finite_differences_of_the_expectation():
    """Linear loss Σ w·L with fixed random weights; for 96 seeds, the adjoint
    gradient at θ and the central difference (θ ± 0.1 on each material's
    roughness, the same seed on both sides): the means of the two agree within
    4.5 standard errors of their difference, per material.  The test has the
    power to see the rejection loop's acceptance term: with that term left out,
    the floor's gradient reads −0.0225 against the difference quotient's
    −0.0007 (standard error 0.0003)."""
    res, rounds, photons = 64, 4, 1 << 16
    base = scenes.adjoint_scene(res, res, seed=1)
    w = np.random.default_rng(5).uniform(0.0, 1.0, size=(res, res, 3)).astype(np.float32)
    h = 0.1
    n_seeds = 96
    grads, fds = [], []
    th = None
    for seed in range(n_seeds):
        cfg = PassConfig(width=res, height=res, photons_per_round=photons, max_photons_per_batch=photons,
                         max_photon_bounces=6, grid_res=32, seed=100 + seed, initial_radius=0.03)
        adj = Adjoint(cfg, base, max_rounds=rounds)
        if th is None:
            th = adj.theta0()
            th[:, 3] = np.clip(th[:, 3], 0.2, 0.85)
        adj.forward(th, rounds)
        grads.append(adj.backward(w / w.size)[:, 3].copy())
        fd = []
        for m in range(th.shape[0] - 1):           # the last material is the light
            tp, tm = th.copy(), th.copy()
            tp[m, 3] += h
            tm[m, 3] -= h
            lp = float((adj.forward(tp, rounds).astype(np.float64) * w).sum() / w.size)
            lm = float((adj.forward(tm, rounds).astype(np.float64) * w).sum() / w.size)
            fd.append((lp - lm) / (2 * h))
        fds.append(fd)
        adj.close()
    g = np.array(grads)[:, :-1]
    f = np.array(fds)
    diff = g - f
    se = diff.std(0, ddof=1) / np.sqrt(n_seeds)
    z = diff.mean(0) / se
    print("adjoint", g.mean(0), "\nfd     ", f.mean(0), "\nz      ", z, flush=True)
    assert np.all(np.abs(z) <= 4.5), (g.mean(0), f.mean(0), se)
    # the roughness gradient as a whole is far from noise
    f_z = f.mean(0) / (f.std(0, ddof=1) / np.sqrt(n_seeds))
    assert np.sum(f_z ** 2) > 50.0, f_z


def test_adjoint_at_specified_size_matches_finite_differences():
    """512², 2^20 photons, 32 passes, 16 bounces, 8 materials: albedo and
    emission gradients of the L2 loss against central differences under
    common random numbers at steps 1e-2 and 3e-3 (2 % per parameter group);
    step 1e-3 is reported as the quotient's noise floor."""
    res, rounds = 512, 32
    cfg = PassConfig(width=res, height=res, photons_per_round=1 << 20, max_photons_per_batch=1 << 20,
                     max_photon_bounces=16, grid_res=64, seed=1)
    scene = scenes.adjoint_scene(res, res, seed=1)
    adj = Adjoint(cfg, scene, max_rounds=rounds)
    th_true = adj.theta0()
    target = adj.forward(th_true, rounds)
    th = th_true.copy()
    th[:, :4] *= np.random.default_rng(7).uniform(0.8, 1.2, size=(th.shape[0], 4)).astype(np.float32)
    img = adj.forward(th, rounds)
    loss, g = l2_loss(img, target)
    grad = adj.backward(g)
    checked = [(m, k) for m in range(th.shape[0]) for k in range(3)] + [(7, 4)]
    for h in (1e-2, 3e-3, 1e-3):
        fd = {}
        for m, k in checked:
            tp, tm = th.copy(), th.copy()
            tp[m, k] += h
            tm[m, k] -= h
            fd[(m, k)] = (l2_loss(adj.forward(tp, rounds), target)[0]
                          - l2_loss(adj.forward(tm, rounds), target)[0]) / (2 * h)
        for k in (0, 1, 2, 4):
            keys = [mk for mk in checked if mk[1] == k]
            a = np.array([grad[m, kk] for m, kk in keys])
            f = np.array([fd[mk] for mk in keys])
            err = np.linalg.norm(a - f) / np.linalg.norm(f)
            print(f"step {h:g} {PARAMS[k]}: relative error {err:.4f}", flush=True)
            if h >= 3e-3:
                assert err <= 0.02, (h, PARAMS[k], a, f)
    adj.close()

