"""Gradients of an SPPM render with respect to per-material parameters.

θ has five entries per material: albedo r, g, b (base_color), GGX roughness
and emission scale (multiplies the material's emissive colour and the power of
the lights built from its triangles; 1 = the scene as given).  ``Adjoint``
wraps phx_adjoint_*: ``forward(θ, rounds)`` renders from a fresh state and
returns the linear estimate; ``backward(dloss_dimage)`` returns ∂loss/∂θ by
replaying the passes with their seeds (detached sampling: the random decisions
are held, the weights are differentiated).  ``render`` exposes the pair as a
``torch.autograd.Function`` on CPU tensors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import PassConfig, check, lib
from .abi import ptr

PARAMS = ("albedo_r", "albedo_g", "albedo_b", "roughness", "emission")


class Adjoint:
    def __init__(self, config: PassConfig, scene, max_rounds: int = 64):
        self.config = config
        self.scene = scene
        self._args = config.to_c()
        self._desc = scene.desc()
        h = C.c_void_p()
        check(lib().phx_adjoint_create(C.byref(self._args), C.byref(self._desc), max_rounds, C.byref(h)),
              "adjoint_create")
        self._h = h
        self.n_materials = int(scene.materials.shape[0])
        self.n_pixels = config.width * config.height

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().phx_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def theta0(self) -> np.ndarray:
        """θ of the scene as given: its albedo and roughness, emission scale 1."""
        m = self.scene.materials
        t = np.zeros((self.n_materials, 5), np.float32)
        t[:, 0:3] = m["base_color"][:, :3]
        t[:, 3] = m["roughness"]
        t[:, 4] = 1.0
        return t

    def forward(self, theta, rounds: int) -> np.ndarray:
        th = np.ascontiguousarray(np.asarray(theta, np.float32).reshape(self.n_materials, 5))
        img = np.zeros((self.n_pixels, 3), np.float32)
        check(lib().phx_adjoint_forward(self._h, ptr(th), rounds, ptr(img)), "adjoint_forward")
        return img.reshape(self.config.height, self.config.width, 3)

    def backward(self, dloss_dimage) -> np.ndarray:
        g = np.ascontiguousarray(np.asarray(dloss_dimage, np.float32).reshape(self.n_pixels, 3))
        out = np.zeros((self.n_materials, 5), np.float32)
        check(lib().phx_adjoint_backward(self._h, ptr(g), ptr(out)), "adjoint_backward")
        return out

    def set_replay_check(self, on: bool) -> None:
        check(lib().phx_adjoint_set_replay_check(self._h, 1 if on else 0), "replay_check")

    def read_replay(self):
        """(emission, direct, deposits per photon, forward deposits per photon) of the replayed first pass."""
        n, nph = self.n_pixels, self.config.photons_per_round
        e, d = np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32)
        c, fc = np.zeros(nph, np.uint32), np.zeros(nph, np.uint32)
        check(lib().phx_adjoint_read_replay(self._h, ptr(e), ptr(d), ptr(c), ptr(fc)), "read_replay")
        return e, d, c, fc

    def replayed_hitpoints(self) -> np.ndarray:
        """Hitpoints of the first pass as the forward trace computes them (after a backward)."""
        r = lib().phx_adjoint_renderer(self._h)
        hp = np.zeros((self.n_pixels, 6, 4), np.float32)
        check(lib().phx_renderer_read_hitpoints(r, ptr(hp)), "read_hitpoints")
        return hp


def l2_loss(image: np.ndarray, target: np.ndarray) -> tuple[float, np.ndarray]:
    """mean((L − T)²) over pixels and channels and its gradient."""
    d = (image.astype(np.float64) - target.astype(np.float64))
    n = d.size
    return float((d * d).sum() / n), (2.0 * d / n).astype(np.float32)


def render(adjoint: Adjoint, rounds: int):
    """A torch.autograd.Function applying ``adjoint`` to a (materials, 5)
    float32 CPU tensor θ: returns the (H, W, 3) estimate; its backward maps
    the image gradient to ∂/∂θ."""
    import torch

    class _Render(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta):
            img = adjoint.forward(theta.detach().cpu().numpy(), rounds)
            return torch.from_numpy(img.copy())

        @staticmethod
        def backward(ctx, grad_image):
            g = adjoint.backward(grad_image.detach().cpu().numpy())
            return torch.from_numpy(g)

    return _Render.apply
