"""phosphor_amd — MI355X (gfx950) SPPM pass loop with the reference's pass interface.

The compute path is the HIP library ``build/libphosphor_hip.so`` (C ABI in
include/phosphor_hip.h).  There is no CPU fallback: loading fails loudly if the
library is missing, and the renderer fails if no HIP device is present.

``Renderer`` mirrors src/renderer/renderer.cpp: construct it with the pass
arguments and a scene (Renderer::Renderer + set_scene), then call
``run_pass()`` per SPPM round or ``run(rounds)`` (Renderer::run).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import abi
from .abi import RenderArgs, ptr

_LIB = None
LIB_PATH = Path(__file__).resolve().parent / "build" / "libphosphor_hip.so"


class PhosphorError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    """The loaded HIP library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        path = Path(os.environ.get("PHOSPHOR_HIP_LIB", LIB_PATH))
        if not path.exists():
            raise RuntimeError(f"{path} is missing: build it with `python -m phosphor_amd.build` "
                               "(there is no CPU fallback for the pass kernels)")
        _LIB = abi.bind(C.CDLL(str(path)))
    return _LIB


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().phx_last_error().decode(errors="replace")
        raise PhosphorError(rc, f"{what}: {msg}" if what else msg)


@dataclass
class PassConfig:
    """The ArgsList fields one SPPM render uses (cmd_args.hpp defaults)."""

    width: int = 1024
    height: int = 1024
    photons_per_round: int = 1 << 20
    max_photons_per_batch: int = 1 << 20
    max_photon_bounces: int = 16
    grid_res: int = 64
    direct_samples: int = 16
    seed: int = 2137
    sppm_alpha: float = 0.7
    initial_radius: float = 0.0
    trace_mode: int = 0  # 0 wavefront trace, 1 one-kernel trace (identical results)
    graph: bool = False  # capture the pass into a hipGraph once and replay it
    emit_wave_bounces: int = 0  # bounce-synchronous emission launches (0 = library default); same results
    reference_bvh: bool = False  # reference median BVH + depth-first order instead of SAH + near-first; same results
    bvh_leaf_size: int = 0  # triangles per leaf of the SAH tree (0 = library default 3); same results
    serial: bool = False  # each gather ends before the next pass starts (no second buffer set); same results
    fused_trace: bool = False  # camera paths in one kernel, photon bounces traced and shaded in one kernel; same results
    camera_block: int = 0  # threads per block of the camera kernel, 64/128/256 (0 = 128); same results
    shadow_block: int = 0  # of the shadow kernel (0 = 128); same results
    emit_block: int = 0  # of the emission bounce kernels (0 = 128); same results
    gather_tile_width: int = 0  # width of the gather's 64-pixel wave tile, 1..64 (0 = 8); same results

    def to_c(self) -> RenderArgs:
        return RenderArgs(self.width, self.height, self.photons_per_round, self.max_photons_per_batch,
                          self.max_photon_bounces, self.grid_res, self.direct_samples, self.seed,
                          self.sppm_alpha, self.initial_radius, self.trace_mode,
                          (abi.PHX_PASS_GRAPH if self.graph else 0) | (abi.PHX_REFERENCE_BVH if self.reference_bvh else 0)
                          | (abi.PHX_PASS_SERIAL if self.serial else 0) | (abi.PHX_FUSED_TRACE if self.fused_trace else 0),
                          self.emit_wave_bounces, self.bvh_leaf_size, self.camera_block, self.shadow_block,
                          self.emit_block, self.gather_tile_width)


def kernel_info() -> dict:
    """Registers and resident waves per SIMD of the pass's main kernels on the
    current device (phx_kernel_info): {name: {vgprs, lds_bytes,
    scratch_bytes, block, waves_per_simd}}."""
    L = lib()
    n = L.phx_kernel_info(None, 0)
    arr = (abi.KernelInfo * max(n, 1))()
    if n >= 0:
        n = L.phx_kernel_info(arr, n)
    if n < 0:
        raise PhosphorError(-n, "phx_kernel_info: " + L.phx_last_error().decode(errors="replace"))
    return {k.name.decode(): {f: int(getattr(k, f)) for f in ("vgprs", "lds_bytes", "scratch_bytes", "block",
                                                                 "waves_per_simd")} for k in arr[:n]}


class Renderer:
    """HIP pass driver (renderer.cpp:9-161) behind the C ABI."""

    def __init__(self, config: PassConfig, scene):
        self.config = config
        self.scene = scene
        self._args = config.to_c()
        self._desc = scene.desc()
        h = C.c_void_p()
        check(lib().phx_renderer_create(C.byref(self._args), C.byref(self._desc), C.byref(h)), "renderer_create")
        self._h = h
        self.n_pixels = config.width * config.height
        self.bucket_count = config.grid_res ** 3 + 1

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().phx_renderer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run_pass(self) -> None:
        check(lib().phx_renderer_run_pass(self._h), "run_pass")

    def run(self, rounds: int) -> None:
        for _ in range(rounds):
            self.run_pass()
        self.synchronize()

    def synchronize(self) -> None:
        check(lib().phx_renderer_synchronize(self._h), "synchronize")

    def stage_ms(self) -> np.ndarray:
        out = np.zeros(4, np.float32)
        check(lib().phx_renderer_last_stage_ms(self._h, ptr(out)), "stage_ms")
        return out

    def gather_ms(self, max_passes: int = 64) -> np.ndarray:
        """The gather kernel's time (ms) in each of the last eager passes since
        reset_gather_ms (at most 64), from HIP events on its stream."""
        out = np.zeros(max_passes, np.float32)
        n = C.c_uint32(0)
        check(lib().phx_renderer_gather_ms(self._h, ptr(out), max_passes, C.byref(n)), "gather_ms")
        return out[:n.value].copy()

    def reset_gather_ms(self) -> None:
        check(lib().phx_renderer_reset_gather_ms(self._h), "reset_gather_ms")

    def pass_ms(self) -> float:
        out = np.zeros(1, np.float32)
        check(lib().phx_renderer_last_pass_ms(self._h, ptr(out)), "pass_ms")
        return float(out[0])

    def checkpoint(self) -> bytes:
        """Progressive state (flux, N, radiance, round count, seed generator)."""
        n = lib().phx_renderer_checkpoint_bytes(self._h)
        buf = np.zeros(n, np.uint8)
        check(lib().phx_renderer_save_checkpoint(self._h, ptr(buf), n), "checkpoint")
        return buf.tobytes()

    def restore(self, data: bytes) -> None:
        buf = np.frombuffer(data, np.uint8).copy()
        check(lib().phx_renderer_load_checkpoint(self._h, ptr(buf), buf.size), "restore")

    def save(self, path) -> None:
        Path(path).write_bytes(self.checkpoint())

    def load(self, path) -> None:
        self.restore(Path(path).read_bytes())

    @property
    def last_photon_count(self) -> int:
        return int(lib().phx_renderer_last_photon_count(self._h))

    @property
    def rounds_done(self) -> int:
        return int(lib().phx_renderer_rounds_done(self._h))

    def read_state(self):
        flux = np.zeros((self.n_pixels, 4), np.float32)
        weight = np.zeros(self.n_pixels, np.float32)
        rad = np.zeros((self.n_pixels, 4), np.float32)
        check(lib().phx_renderer_read_state(self._h, ptr(flux), ptr(weight), ptr(rad)), "read_state")
        return flux, weight, rad

    def read_hitpoints(self) -> np.ndarray:
        hp = np.zeros((self.n_pixels, 6, 4), np.float32)
        check(lib().phx_renderer_read_hitpoints(self._h, ptr(hp)), "read_hitpoints")
        return hp

    def read_photon_map(self):
        n = min(self.last_photon_count, self.config.photons_per_round * self.config.max_photon_bounces)
        pos = np.zeros((n, 4), np.float32)
        power = np.zeros(n, np.uint32)
        dirs = np.zeros(n, np.uint32)
        cs = np.zeros(self.bucket_count, np.uint32)
        ce = np.zeros(self.bucket_count, np.uint32)
        check(lib().phx_renderer_read_photon_map(self._h, ptr(pos), ptr(power), ptr(dirs), ptr(cs), ptr(ce)),
              "read_photon_map")
        return pos, power, dirs, cs, ce

    def estimate(self) -> np.ndarray:
        """Linear radiance estimate (output.cpp:52-83 before tonemapping), (H, W, 3)."""
        rgb = np.zeros((self.n_pixels, 3), np.float32)
        check(lib().phx_renderer_estimate(self._h, ptr(rgb)), "estimate")
        return rgb.reshape(self.config.height, self.config.width, 3)

    def device_buffers(self) -> list[int]:
        out = (C.c_void_p * 20)()
        check(lib().phx_renderer_device_buffers(self._h, out, 20), "device_buffers")
        return [int(v or 0) for v in out]


class Batch:
    """Many scenes rendered together (phx_batch_*): one pass over every scene,
    each scene's results bit-identical to a ``Renderer`` with its config."""

    def __init__(self, configs: list[PassConfig], scenes: list, trace_group_pixels: int = 0,
                 emit_group_photons: int = 0, graph: bool = True, serial: bool = False):
        if len(configs) != len(scenes) or not scenes:
            raise ValueError("one PassConfig per scene, at least one scene")
        self.configs = list(configs)
        self.scenes = list(scenes)
        n = len(scenes)
        self._args = (RenderArgs * n)(*(c.to_c() for c in configs))
        self._descs = [sc.desc() for sc in scenes]  # keeps the host arrays referenced
        self._desc_arr = (abi.SceneDesc * n)(*self._descs)
        flags = (abi.PHX_PASS_GRAPH if graph else 0) | (abi.PHX_PASS_SERIAL if serial else 0)
        opt = abi.BatchOptions(trace_group_pixels, emit_group_photons, flags, 0)
        h = C.c_void_p()
        check(lib().phx_batch_create(C.cast(self._args, C.c_void_p), C.cast(self._desc_arr, C.c_void_p), n,
                                     C.byref(opt), C.byref(h)), "batch_create")
        self._h = h
        self.n_scenes = n

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().phx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_pass(self) -> None:
        check(lib().phx_batch_run_pass(self._h), "batch_run_pass")

    def run(self, rounds: int) -> None:
        for _ in range(rounds):
            self.run_pass()
        self.synchronize()

    def synchronize(self) -> None:
        check(lib().phx_batch_synchronize(self._h), "batch_synchronize")

    def gather_ms(self, max_passes: int = 64) -> np.ndarray:
        """The gather kernel's time (ms) in each of the last eager passes since
        reset_gather_ms (at most 64), from HIP events on its stream."""
        out = np.zeros(max_passes, np.float32)
        n = C.c_uint32(0)
        check(lib().phx_renderer_gather_ms(self._h, ptr(out), max_passes, C.byref(n)), "gather_ms")
        return out[:n.value].copy()

    def reset_gather_ms(self) -> None:
        check(lib().phx_renderer_reset_gather_ms(self._h), "reset_gather_ms")

    def pass_ms(self) -> float:
        out = np.zeros(1, np.float32)
        check(lib().phx_batch_last_pass_ms(self._h, ptr(out)), "batch_pass_ms")
        return float(out[0])

    @property
    def rounds_done(self) -> int:
        return int(lib().phx_batch_rounds_done(self._h))

    def photon_counts(self) -> np.ndarray:
        out = np.zeros(self.n_scenes, np.uint32)
        check(lib().phx_batch_photon_counts(self._h, ptr(out)), "batch_photon_counts")
        return out

    def _npx(self, s: int) -> int:
        return self.configs[s].width * self.configs[s].height

    def read_state(self, s: int):
        n = self._npx(s)
        flux, weight, rad = np.zeros((n, 4), np.float32), np.zeros(n, np.float32), np.zeros((n, 4), np.float32)
        check(lib().phx_batch_read_state(self._h, s, ptr(flux), ptr(weight), ptr(rad)), "batch_read_state")
        return flux, weight, rad

    def read_hitpoints(self, s: int) -> np.ndarray:
        hp = np.zeros((self._npx(s), 6, 4), np.float32)
        check(lib().phx_batch_read_hitpoints(self._h, s, ptr(hp)), "batch_read_hitpoints")
        return hp

    def read_photon_map(self, s: int):
        n = int(self.photon_counts()[s])
        nb = self.configs[s].grid_res ** 3 + 1
        pos, power, dirs = np.zeros((n, 4), np.float32), np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        cs, ce = np.zeros(nb, np.uint32), np.zeros(nb, np.uint32)
        check(lib().phx_batch_read_photon_map(self._h, s, ptr(pos), ptr(power), ptr(dirs), ptr(cs), ptr(ce)),
              "batch_read_photon_map")
        return pos, power, dirs, cs, ce

    def estimate(self, s: int) -> np.ndarray:
        c = self.configs[s]
        rgb = np.zeros((c.width * c.height, 3), np.float32)
        check(lib().phx_batch_estimate(self._h, s, ptr(rgb)), "batch_estimate")
        return rgb.reshape(c.height, c.width, 3)


def config_for(name: str, **overrides) -> tuple[PassConfig, object, int]:
    """PassConfig, scene and round count of a named workload (phosphor_amd.scenes.CONFIGS)."""
    from . import scenes

    spec = dict(scenes.CONFIGS[name])
    spec.update(overrides)
    scene = spec["scene"](spec["width"], spec["height"])
    lo, hi = scene.bounds()
    extent = float((hi - lo).max()) + 2e-3
    radius = spec.get("radius")
    grid = spec.get("grid_res") or (scenes.grid_res_for_radius(extent, radius) if radius else 64)
    cfg = PassConfig(width=spec["width"], height=spec["height"], photons_per_round=spec["photons_per_round"],
                     max_photons_per_batch=min(spec["photons_per_round"], 1 << 20),
                     max_photon_bounces=spec["max_photon_bounces"], grid_res=grid,
                     direct_samples=spec.get("direct_samples", 16), seed=spec["seed"],
                     sppm_alpha=spec["sppm_alpha"], initial_radius=float(radius or 0.0))
    return cfg, scene, spec["rounds"]


__all__ = ["Renderer", "Batch", "PassConfig", "PhosphorError", "lib", "check", "config_for", "abi"]
