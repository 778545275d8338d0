"""The reference renderer's pass through the stage-level C ABI.

``StagePass`` mirrors ``Renderer::set_scene`` / ``Renderer::run`` of
src/renderer/renderer.cpp:22-161 the way a maintainer would wire the drop-in
(the sequence INTEGRATION.md shows in C++): the host owns every buffer, the
round seed comes from the host PCG generator (renderer.cpp:54,65), and each
pass enqueues, on one stream and in the reference's order,

  trace_rays → emit_photons (batches, then the blocking counter readback of
  renderer.cpp:113-118) → hash_photons → scan_cells → sort_photons →
  gather_flux,

calling ``phx_trace_rays_wavefront``, ``phx_emit_photons``,
``phx_hash_photons``, ``phx_scan_cells``, ``phx_sort_photons`` and
``phx_gather_flux``.  Traversal uses the reference's median BVH (the
kernels' depth-first search of it), the photon map is ordered by cell with
deposits in canonical order inside a cell.  HIP events between the stages
give per-stage times; the library's own pass driver (``Renderer``) is the
fast path, this one measures the drop-in boundary.

``fast=True`` wires the drop-in's faster forms of the same stages:
``phx_accel_attach`` once at set-up (the trace and emission entry points then
trace the library's SAH child-pair tree, same hits), and
``phx_sort_photons_by_octant`` + ``phx_gather_flux_octants`` in place of
hash → scan → sort → gather (the map ordered by cell, then half-cell octant,
records decoded once; stages "trace", "emit", "map", "gather").  The results
equal the oracle's pass with an octant-ordered map, bit for bit.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import PassConfig, check, lib
from .abi import HASH_DTYPE

M32 = (1 << 32) - 1
STAGES = ("trace", "emit", "hash", "scan", "sort", "gather")
FAST_STAGES = ("trace", "emit", "map", "gather")


def _pcg_seed(seed: int) -> int:
    """random.h:16-25: state 0, step, add the seed, step."""
    s = (0 * 747796405 + 2891336453) & M32
    s = (s + seed) & M32
    return (s * 747796405 + 2891336453) & M32


def _pcg_next(state: int) -> tuple[int, int]:
    """random.h:27-34: step the LCG, RXS-M-XS output of the old state."""
    word = ((state >> ((state >> 28) + 4)) ^ state) & M32
    return (state * 747796405 + 2891336453) & M32, (word * 277803737) & M32


class _Hip:
    """HIP events of the runtime the library runs on."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so")
        self.rt.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.rt.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.rt.hipEventSynchronize.argtypes = [C.c_void_p]
        self.rt.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.rt.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        if self.rt.hipEventCreate(C.byref(e)) != 0:
            raise RuntimeError("hipEventCreate failed")
        return e

    def record(self, e):
        if self.rt.hipEventRecord(e, None) != 0:
            raise RuntimeError("hipEventRecord failed")

    def elapsed(self, a, b) -> float:
        if self.rt.hipEventSynchronize(b) != 0:
            raise RuntimeError("hipEventSynchronize failed")
        ms = C.c_float()
        if self.rt.hipEventElapsedTime(C.byref(ms), a, b) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return float(ms.value)


class _Buf:
    """One device allocation of the library's runtime."""

    def __init__(self, nbytes: int, host: np.ndarray | None = None, fill: int | None = 0):
        self.nbytes = max(int(nbytes), 16)
        p = C.c_void_p()
        check(lib().phx_device_malloc(C.byref(p), self.nbytes), "device_malloc")
        self.ptr = int(p.value)
        if host is not None:
            h = np.ascontiguousarray(host)
            if h.nbytes:
                check(lib().phx_memcpy_htod(self.ptr, h.ctypes.data, h.nbytes), "upload")
        elif fill is not None:
            check(lib().phx_memset_device(self.ptr, fill, self.nbytes), "clear")

    def read(self, dtype, shape) -> np.ndarray:
        out = np.zeros(shape, dtype)
        if out.nbytes:
            check(lib().phx_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes), "download")
        return out

    def free(self):
        if self.ptr:
            lib().phx_device_free(self.ptr)
            self.ptr = 0


class StagePass:
    """Renderer::set_scene + Renderer::run over the stage-level C ABI."""

    def __init__(self, config: PassConfig, scene, fast: bool = False):
        L = lib()
        self.fast = fast
        self.stages = FAST_STAGES if fast else STAGES
        c = self.config = config
        self.scene = sc = scene
        self.n_pixels = c.width * c.height
        self.photon_buf = c.photons_per_round * c.max_photon_bounces       # renderer.cpp:40
        self.bucket_count = c.grid_res ** 3 + 1                               # renderer.cpp:38
        bufs = []

        def buf(*a, **k):
            b = _Buf(*a, **k)
            bufs.append(b)
            return b

        self._bufs = bufs
        # scene (geometry.cpp / shading.cpp): the reference's median BVH
        n = sc.n_triangles
        nodes = np.zeros((2 * n - 1, 8), np.float32)
        if L.phx_build_bvh_median(sc.tri_verts.ctypes.data, n, nodes.ctypes.data) != 2 * n - 1:
            raise RuntimeError("phx_build_bvh_median failed")
        self.hash = np.zeros(1, HASH_DTYPE)
        check(L.phx_build_photon_hash_info(nodes[:1].ctypes.data, c.grid_res, self.hash.ctypes.data), "hash info")
        lo, hi = nodes[0, :3].astype(np.float32), nodes[0, 4:7].astype(np.float32)
        center = (np.float32(0.5) * (lo + hi)).astype(np.float32)             # shading.cpp:45-58
        radius = np.float32(0.0)
        for k in range(8):
            d = np.array([hi[0] if k & 1 else lo[0], hi[1] if k & 2 else lo[1], hi[2] if k & 4 else lo[2]],
                         np.float32) - center
            radius = max(radius, np.float32(math.sqrt(float(np.float32(np.float32(d[0] * d[0] + d[1] * d[1])
                                                                         + d[2] * d[2])))))
        self.center = np.array([*center, 0.0], np.float32)
        self.radius = float(radius)
        lum, run = [], np.float32(0.0)                                        # scene_data.cpp:14-23
        for p in sc.lights["power"]:
            v = np.float32(np.float32(p[0] * np.float32(0.2126)) + np.float32(p[1] * np.float32(0.7152)))
            v = np.float32(v + np.float32(p[2] * np.float32(0.0722)))
            run = np.float32(run + max(v, np.float32(0.0)))
            lum.append(run)
        self.total_lum = float(lum[-1]) if lum else 0.0
        self.tv, self.tn, self.tt, self.tuv, self.tmat = (buf(a.nbytes, a) for a in (
            sc.tri_verts, sc.tri_normals, sc.tri_tangents, sc.tri_uvs, sc.tri_mat_index))
        self.nodes = buf(nodes.nbytes, nodes)
        self.accelerated = False
        if fast:
            acc = C.c_uint32(0)
            check(L.phx_accel_attach(self.nodes.ptr, self.tv.ptr, nodes.ctypes.data, nodes.shape[0],
                                     sc.tri_verts.ctypes.data, n, 0, C.byref(acc)), "accel_attach")
            self.accelerated = bool(acc.value)
        self.etv, self.etn, self.etuv, self.etmat = (buf(a.nbytes, a) for a in (
            sc.etri_verts, sc.etri_normals, sc.etri_uvs, sc.etri_mat_index))
        self.mats, self.uvt, self.lights = (buf(a.nbytes, a) for a in (sc.materials, sc.uv_transforms, sc.lights))
        self.lum = buf(4 * max(len(lum), 1), np.array(lum, np.float32))
        self.tex_meta = buf(max(sc.tex_meta.nbytes, 16), sc.tex_meta if sc.tex_meta.nbytes else None)
        self.atlas = buf(max(sc.tex_atlas.nbytes, 16), sc.tex_atlas)
        env = sc.envmap
        self.env = buf(env.nbytes, env) if env is not None else None
        # image buffers (image.cpp:11-21), photon buffers, photon map (photon.cpp)
        npx = self.n_pixels
        self.hp = [buf(16 * npx) for _ in range(6)]
        flux = np.zeros((npx, 4), np.float32)
        if c.initial_radius > 0:
            flux[:, 3] = np.float32(c.initial_radius) ** 2
        self.flux = buf(flux.nbytes, flux)
        self.weight = buf(4 * npx)
        self.radiance = buf(16 * npx)
        pb = self.photon_buf
        self.pos, self.pow, self.dir = buf(16 * pb), buf(4 * pb), buf(4 * pb)
        self.counter = buf(4)
        self.hashes, self.counts = buf(4 * pb), buf(4 * self.bucket_count)
        self.cstart, self.cend = buf(4 * self.bucket_count), buf(4 * self.bucket_count)
        self.spos, self.spow, self.sdir = buf(16 * pb), buf(4 * pb), buf(4 * pb)
        self.per_batch = min(c.photons_per_round, c.max_photons_per_batch or 1)
        self.trace_wsb = L.phx_trace_rays_wavefront_workspace_size(c.width, c.height, c.direct_samples)
        self.emit_wsb = L.phx_emit_photons_workspace_size(self.per_batch, c.max_photon_bounces)
        self.scan_wsb = L.phx_scan_cells_workspace_size(self.bucket_count)
        self.sort_wsb = L.phx_sort_photons_workspace_size(pb, self.bucket_count)
        self.trace_ws, self.emit_ws = buf(self.trace_wsb, fill=None), buf(self.emit_wsb, fill=None)
        self.scan_ws, self.sort_ws = buf(self.scan_wsb, fill=None), buf(self.sort_wsb, fill=None)
        if fast:
            self.oct_start = buf(32 * self.bucket_count)
            self.dirpow = buf(16 * pb)
            self.map_wsb = L.phx_sort_photons_by_octant_workspace_size(pb, self.bucket_count)
            self.map_ws = buf(self.map_wsb, fill=None)
        self.rng = _pcg_seed(c.seed)                                          # renderer.cpp:54
        self.rounds = 0
        self.total_photons = 0
        self.last_photon_count = 0
        self.hip = _Hip()
        self.events = [self.hip.event() for _ in range(len(self.stages) + 1)]
        self.stage_ms = np.zeros(len(self.stages))

    def close(self):
        if getattr(self, "accelerated", False):
            lib().phx_accel_detach(self.nodes.ptr)
            self.accelerated = False
        for b in getattr(self, "_bufs", []):
            b.free()
        self._bufs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_pass(self) -> None:
        """renderer.cpp:64-78: one round."""
        L, c, sc = lib(), self.config, self.scene
        self.rng, seed = _pcg_next(self.rng)                                  # renderer.cpp:65
        ev, hip = self.events, self.hip
        hip.record(ev[0])
        env = self.env
        check(L.phx_trace_rays_wavefront(                                     # renderer.cpp:90-95
            sc.camera.ctypes.data, c.width, c.height, seed, c.direct_samples, self.tv.ptr, self.tn.ptr, self.tt.ptr,
            self.tuv.ptr, self.nodes.ptr, self.tmat.ptr, sc.n_triangles, self.mats.ptr, self.uvt.ptr,
            self.tex_meta.ptr, self.atlas.ptr, self.lights.ptr, len(sc.lights), self.lum.ptr, self.total_lum,
            self.center.ctypes.data, self.radius, self.etv.ptr, self.etn.ptr, self.etuv.ptr, self.etmat.ptr,
            env.ptr if env else None, sc.envmap.shape[1] if env else 0, sc.envmap.shape[0] if env else 0,
            1 if env else 0, *(b.ptr for b in self.hp), 0, self.trace_ws.ptr, self.trace_wsb, None), "trace_rays")
        hip.record(ev[1])
        check(L.phx_memset_device(self.counter.ptr, 0, 4), "counter")       # renderer.cpp:99
        for off in range(0, c.photons_per_round, self.per_batch):            # renderer.cpp:103-111
            cnt = min(self.per_batch, c.photons_per_round - off)
            check(L.phx_emit_photons(
                self.lights.ptr, len(sc.lights), self.photon_buf, off, cnt, c.max_photon_bounces, seed,
                self.nodes.ptr, self.tv.ptr, self.tn.ptr, self.tt.ptr, self.tuv.ptr, self.tmat.ptr, sc.n_triangles,
                self.etv.ptr, self.etn.ptr, self.etuv.ptr, self.etmat.ptr, self.mats.ptr, self.uvt.ptr,
                self.tex_meta.ptr, self.atlas.ptr, self.lum.ptr, self.total_lum, self.center.ctypes.data,
                self.radius, self.pos.ptr, self.pow.ptr, self.dir.ptr, self.counter.ptr, self.emit_ws.ptr,
                self.emit_wsb, None), "emit_photons")
        count = int(self.counter.read(np.uint32, (1,))[0])                   # renderer.cpp:113-118
        n = min(count, self.photon_buf)
        self.last_photon_count = count
        if n == 0:
            raise RuntimeError("no photon hit")
        hip.record(ev[2])
        if self.fast:
            check(L.phx_sort_photons_by_octant(self.pos.ptr, self.pow.ptr, self.dir.ptr, n, self.hash.ctypes.data,
                                               self.bucket_count, self.spos.ptr, self.spow.ptr, self.sdir.ptr,
                                               self.dirpow.ptr, self.cstart.ptr, self.cend.ptr, self.oct_start.ptr,
                                               self.map_ws.ptr, self.map_wsb, None), "sort_photons_by_octant")
            hip.record(ev[3])
            check(L.phx_gather_flux_octants(*(b.ptr for b in self.hp), c.width, c.height, self.spos.ptr,
                                            self.dirpow.ptr, n, self.cstart.ptr, self.cend.ptr, self.oct_start.ptr,
                                            self.hash.ctypes.data, c.sppm_alpha, self.flux.ptr, self.weight.ptr,
                                            self.radiance.ptr, None), "gather_flux_octants")
            hip.record(ev[4])
            self._finish(n)
            return
        check(L.phx_memset_device(self.counts.ptr, 0, 4 * self.bucket_count), "counts")   # renderer.cpp:128
        check(L.phx_hash_photons(self.pos.ptr, n, self.hash.ctypes.data, self.hashes.ptr, self.counts.ptr, None),
              "hash_photons")
        hip.record(ev[3])
        check(L.phx_scan_cells(self.counts.ptr, self.bucket_count, self.cstart.ptr, self.cend.ptr, self.scan_ws.ptr,
                               self.scan_wsb, None), "scan_cells")
        hip.record(ev[4])
        check(L.phx_sort_photons(self.pos.ptr, self.pow.ptr, self.dir.ptr, self.hashes.ptr, self.cstart.ptr,
                                 self.cend.ptr, n, self.bucket_count, self.spos.ptr, self.spow.ptr, self.sdir.ptr,
                                 self.sort_ws.ptr, self.sort_wsb, None), "sort_photons")
        hip.record(ev[5])
        check(L.phx_gather_flux(*(b.ptr for b in self.hp), self.n_pixels, self.spos.ptr, self.spow.ptr,
                                self.sdir.ptr, n, self.cstart.ptr, self.cend.ptr, self.hash.ctypes.data,
                                c.sppm_alpha, self.flux.ptr, self.weight.ptr, self.radiance.ptr, None),
              "gather_flux")
        hip.record(ev[6])
        self._finish(n)

    def _finish(self, n: int) -> None:
        ev, hip = self.events, self.hip
        self.stage_ms = np.array([hip.elapsed(ev[k], ev[k + 1]) for k in range(len(self.stages))])
        self.last_n = n
        self.rounds += 1
        self.total_photons += self.config.photons_per_round                    # renderer.cpp:78

    def read_state(self):
        npx = self.n_pixels
        return (self.flux.read(np.float32, (npx, 4)), self.weight.read(np.float32, (npx,)),
                self.radiance.read(np.float32, (npx, 4)))

    def read_photon_map(self):
        n = self.last_n
        return (self.spos.read(np.float32, (n, 4)), self.spow.read(np.uint32, (n,)), self.sdir.read(np.uint32, (n,)),
                self.cstart.read(np.uint32, (self.bucket_count,)), self.cend.read(np.uint32, (self.bucket_count,)))
