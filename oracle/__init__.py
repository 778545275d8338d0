"""CPU oracle of the SPPM pass (TEST INFRASTRUCTURE ONLY).

Python binding of ``oracle/build/librefcpu.so`` (built from refcpu.cpp by the
Makefile beside it).  Only the tests, the smoke check and the benchmark's
CPU-baseline leg use this package; the product (phosphor_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
# PHOSPHOR_ORACLE_LIB selects another build of refcpu.cpp (the planted-mistake
# test loads mutated copies this way)
LIB_PATH = Path(os.environ.get("PHOSPHOR_ORACLE_LIB", ORACLE_DIR / "build" / "librefcpu.so"))
_LIB = None

VP, U32, F32, INT = C.c_void_p, C.c_uint32, C.c_float, C.c_int

_SIGS = {
    "ref_num_threads": (INT,),
    "ref_set_num_threads": (None, INT),
    "ref_pcg_seed": (U32, U32),
    "ref_random_u32": (U32, VP),
    "ref_random_float": (F32, VP),
    "ref_encode_rgbe": (U32, VP),
    "ref_decode_rgbe": (None, U32, VP),
    "ref_encode_oct": (U32, VP),
    "ref_decode_oct": (None, U32, VP),
    "ref_get_hash": (U32, VP, VP),
    "ref_math": (None, INT, VP, VP, VP, U32),
    "ref_build_bvh": (U32, VP, U32, VP),
    "ref_scene_intersect": (INT, VP, VP, VP, U32, VP, VP, F32, F32, VP),
    "ref_sample_bsdf": (U32, VP, VP, VP, VP, INT, VP, VP),
    "ref_sample_bsdf_ior": (U32, VP, VP, VP, VP, INT, VP, VP, VP, VP),
    "ref_renderer_camera_ray": (U32, VP, U32, U32, VP, VP),
    "ref_emit_photons_counted": (U32, VP, U32, U32, U32, VP, VP, VP, VP, U32),
    "ref_build_photon_hash_info": (None, VP, U32, VP),
    "ref_hash_photons": (None, VP, U32, VP, VP, VP),
    "ref_scan_cells": (None, VP, U32, VP, VP),
    "ref_sort_photons": (None, VP, VP, VP, VP, VP, U32, VP, VP, VP),
    "ref_gather_flux": (None, VP, VP, VP, VP, VP, VP, U32, VP, VP, VP, VP, VP, VP, F32, VP, VP, VP),
    "ref_renderer_create": (VP, VP, VP),
    "ref_renderer_destroy": (None, VP),
    "ref_renderer_count_visits": (None, VP, INT),
    "ref_renderer_octant_order": (None, VP, INT),
    "ref_renderer_rays": (None, VP, VP),
    "ref_trace_range": (None, VP, U32, U32, U32),
    "ref_emit_range": (U32, VP, U32, U32, U32, VP, VP, VP, U32),
    "ref_renderer_run_pass": (INT, VP),
    "ref_renderer_next_seed": (U32, VP),
    "ref_renderer_sample_light": (U32, VP, U32, U32, VP, VP, VP),
    "ref_renderer_stage_ms": (None, VP, VP),
    "ref_renderer_visits": (None, VP, VP),
    "ref_renderer_last_photon_count": (U32, VP),
    "ref_renderer_rounds_done": (U32, VP),
    "ref_renderer_hash_info": (None, VP, VP),
    "ref_renderer_bvh_nodes": (U32, VP, VP, U32),
    "ref_renderer_scene_sphere": (None, VP, VP, VP),
    "ref_renderer_read_state": (None, VP, VP, VP, VP),
    "ref_renderer_write_state": (None, VP, VP, VP, VP),
    "ref_renderer_read_hitpoints": (None, VP, VP),
    "ref_renderer_read_deposits": (U32, VP, VP, VP, VP),
    "ref_renderer_read_photon_map": (U32, VP, VP, VP, VP, VP, VP),
    "ref_renderer_estimate": (None, VP, VP),
}


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, *args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = list(args)
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data


class RefRenderer:
    """CPU restatement of Renderer (renderer.cpp) with the same pass interface."""

    def __init__(self, config, scene, count_visits: bool = False, octant_order: bool = True):
        from phosphor_amd.abi import HASH_DTYPE  # layouts only

        self.config = config
        self.scene = scene
        self._args = config.to_c()
        self._desc = scene.desc()
        self._h = lib().ref_renderer_create(C.byref(self._args), C.byref(self._desc))
        if not self._h:
            raise RuntimeError("ref_renderer_create failed")
        lib().ref_renderer_count_visits(self._h, 1 if count_visits else 0)
        lib().ref_renderer_octant_order(self._h, 1 if octant_order else 0)
        self.n_pixels = config.width * config.height
        self.bucket_count = config.grid_res ** 3 + 1
        self._hash_dtype = HASH_DTYPE

    def close(self):
        if getattr(self, "_h", None):
            lib().ref_renderer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_pass(self) -> int:
        return int(lib().ref_renderer_run_pass(self._h))

    def next_seed(self) -> int:
        return int(lib().ref_renderer_next_seed(self._h))

    def trace_range(self, seed: int, first: int, count: int) -> None:
        lib().ref_trace_range(self._h, seed, first, count)

    def emit_range(self, seed: int, first: int, count: int) -> int:
        return int(lib().ref_emit_range(self._h, seed, first, count, None, None, None, 0))

    def camera_ray(self, seed: int, tid: int):
        """Origin and direction of pixel `tid`'s camera ray in a pass with `seed`."""
        o, d = np.zeros(4, np.float32), np.zeros(4, np.float32)
        lib().ref_renderer_camera_ray(self._h, seed, tid, _p(o), _p(d))
        return o, d

    def emit_counted(self, seed: int, first: int, count: int):
        """Deposits of photons [first, first+count) in (photon, bounce) order and
        the number each photon made: (counts, pos, power, dir)."""
        counts = np.zeros(count, np.uint32)
        cap = count * self.config.max_photon_bounces
        pos = np.zeros((cap, 4), np.float32)
        pw = np.zeros(cap, np.uint32)
        dr = np.zeros(cap, np.uint32)
        n = lib().ref_emit_photons_counted(self._h, seed, first, count, _p(counts), _p(pos), _p(pw), _p(dr), cap)
        return counts, pos[:n], pw[:n], dr[:n]

    def sample_light(self, seed: int, tid: int):
        o, d, p = (np.zeros(4, np.float32) for _ in range(3))
        lib().ref_renderer_sample_light(self._h, seed, tid, _p(o), _p(d), _p(p))
        return o, d, p

    def stage_ms(self) -> np.ndarray:
        out = np.zeros(4, np.float64)
        lib().ref_renderer_stage_ms(self._h, _p(out))
        return out

    def visits(self) -> np.ndarray:
        out = np.zeros(3, np.uint64)
        lib().ref_renderer_visits(self._h, _p(out))
        return out

    def rays(self) -> np.ndarray:
        """scene_intersect calls of the trace and emission stages so far
        (counted only with count_visits=True)."""
        out = np.zeros(2, np.uint64)
        lib().ref_renderer_rays(self._h, _p(out))
        return out

    @property
    def last_photon_count(self) -> int:
        return int(lib().ref_renderer_last_photon_count(self._h))

    def hash_info(self) -> np.ndarray:
        out = np.zeros(1, self._hash_dtype)
        lib().ref_renderer_hash_info(self._h, _p(out))
        return out

    def bvh_nodes(self) -> np.ndarray:
        n = lib().ref_renderer_bvh_nodes(self._h, None, 0)
        out = np.zeros((n, 8), np.float32)
        lib().ref_renderer_bvh_nodes(self._h, _p(out), n)
        return out

    def scene_sphere(self):
        c = np.zeros(4, np.float32)
        r = np.zeros(1, np.float32)
        lib().ref_renderer_scene_sphere(self._h, _p(c), _p(r))
        return c, float(r[0])

    def read_state(self):
        flux = np.zeros((self.n_pixels, 4), np.float32)
        weight = np.zeros(self.n_pixels, np.float32)
        rad = np.zeros((self.n_pixels, 4), np.float32)
        lib().ref_renderer_read_state(self._h, _p(flux), _p(weight), _p(rad))
        return flux, weight, rad

    def write_state(self, flux, weight, rad):
        lib().ref_renderer_write_state(self._h, _p(np.ascontiguousarray(flux, np.float32)),
                                       _p(np.ascontiguousarray(weight, np.float32)),
                                       _p(np.ascontiguousarray(rad, np.float32)))

    def read_hitpoints(self) -> np.ndarray:
        hp = np.zeros((self.n_pixels, 6, 4), np.float32)
        lib().ref_renderer_read_hitpoints(self._h, _p(hp))
        return hp

    def read_deposits(self):
        n = lib().ref_renderer_read_deposits(self._h, None, None, None)
        pos = np.zeros((n, 4), np.float32)
        pw = np.zeros(n, np.uint32)
        dr = np.zeros(n, np.uint32)
        lib().ref_renderer_read_deposits(self._h, _p(pos), _p(pw), _p(dr))
        return pos, pw, dr

    def read_photon_map(self):
        n = lib().ref_renderer_read_photon_map(self._h, None, None, None, None, None)
        pos = np.zeros((n, 4), np.float32)
        pw = np.zeros(n, np.uint32)
        dr = np.zeros(n, np.uint32)
        cs = np.zeros(self.bucket_count, np.uint32)
        ce = np.zeros(self.bucket_count, np.uint32)
        lib().ref_renderer_read_photon_map(self._h, _p(pos), _p(pw), _p(dr), _p(cs), _p(ce))
        return pos, pw, dr, cs, ce

    def estimate(self) -> np.ndarray:
        rgb = np.zeros((self.n_pixels, 3), np.float32)
        lib().ref_renderer_estimate(self._h, _p(rgb))
        return rgb.reshape(self.config.height, self.config.width, 3)
