"""GPU parity on two rules no workload reaches: the traversal stack's overflow
rule (hit.h:163-164) and the thin-lens camera (camera.h:28-47).

Stack overflow.  The reference's depth-first search keeps a 24-entry stack
and, when a pop leaves its pointer p with p + 2 >= 24, drops both children of
the interior node it is visiting.  A left-deep comb of 30 interior levels,
handed to the renderer and to phx_trace_rays as a prebuilt reference tree,
makes that happen: interior node j is visited with p = j - 1, so the right
leaves of nodes 0..22 stay reachable and everything below node 23 is lost.
The scene puts a floor in leaf 0 and a panel facing the camera in each leaf
k = 1..30 (at x = k); camera rays, shadow rays and photons all traverse the
comb.  Both GPU paths must equal the oracle given the same tree, bit for bit,
and the oracle's image must show panels 1..22 and none of 23..30 (with its
own median tree it shows all 30), so the rule is exercised, not bypassed.

Thin lens.  The caustic scene with a defocused camera (aperture angle 1.5°,
focus at 2.0): two whole passes bit-exact against the oracle.
"""
import dataclasses

import numpy as np
import pytest

import phosphor_amd as P
from phosphor_amd import PassConfig, Renderer, scenes

import oracle

from tests.gpu_util import dptr, empty_dev, from_dev
from tests.test_gpu_kernels_abi import DeviceScene, _check_wavefront
from tests.test_gpu_pass import _bits, _compare_pass

pytestmark = pytest.mark.gpu

DEPTH = 30
INTERIOR = np.uint32(0xFFFFFFFF).view(np.float32)


def _comb_nodes(tris):
    """Reference-layout nodes (min.w = triangle or 0xFFFFFFFF for interior,
    max.w = right child) of a left-deep comb: interior node j has leaf j as
    its right child; the deepest interior node holds leaves DEPTH and
    DEPTH - 1.  Interior boxes span the whole scene."""
    lo_all, hi_all = [-3.0, -1.0, -4.0], [2 * DEPTH + 14.0, 3.0, 10.0]
    nodes = []

    def leaf(k):
        lo = tris[k].min(0) - np.float32(1e-3)
        hi = tris[k].max(0) + np.float32(1e-3)
        nodes.append([*lo, np.uint32(k).view(np.float32), *hi, np.float32(0)])

    def interior(j):
        at = len(nodes)
        nodes.append(None)
        if j + 1 < DEPTH:
            interior(j + 1)
        else:
            leaf(DEPTH)
        right = len(nodes)
        leaf(j)
        nodes[at] = [*lo_all, INTERIOR, *hi_all, np.uint32(right).view(np.float32)]

    interior(0)
    out = np.zeros((len(nodes), 8), np.float32)
    for i, n in enumerate(nodes):
        out[i] = np.array(n, np.float32)
        out[i, 3] = n[3]           # keep the exact bit patterns of the w lanes
        out[i, 7] = n[7]
    return out


def _comb_scene(width, height, reference_tree=True):
    b = scenes.SceneBuilder()
    m = b.material(scenes.default_material((0.7, 0.7, 0.7)))
    tris = np.zeros((DEPTH + 1, 3, 3), np.float32)
    tris[0] = [(-2, 0, -3), (-2, 0, 9), (2 * DEPTH + 12, 0, -3)]       # floor, facing +y
    for k in range(1, DEPTH + 1):
        tris[k] = [(k, 0.1, 0), (k, 1.0, 0), (k + 0.8, 0.1, 0)]         # panel at x = k, facing -z
    b.add_mesh(tris, m)
    b.add_point_light((DEPTH / 2, 4.0, -3.0), (300.0, 300.0, 300.0))
    cx = DEPTH / 2 + 0.5
    cam = scenes.make_camera((cx, 0.55, -12.0), (cx, 0.55, 0.0), (0.0, 1.0, 0.0), 108.0, width / height)
    sc = b.build(cam, name="comb")
    if reference_tree:
        sc = dataclasses.replace(sc, bvh_nodes=_comb_nodes(tris))
    return sc


def _panels_seen(hp):
    """x of the camera hits on the panels (diffuse hits mark emission.w = 1)."""
    valid = hp[:, 5, 3] == 1.0
    pos = hp[:, 0, :3]
    on_panel = valid & (np.abs(pos[:, 2]) < 1e-3) & (pos[:, 1] > 0.1)
    return np.unique(np.floor(pos[on_panel, 0]).astype(int))


def _config():
    return PassConfig(width=160, height=40, photons_per_round=1 << 14, max_photons_per_batch=1 << 14,
                      max_photon_bounces=6, grid_res=40, seed=5, initial_radius=0.05)


def test_comb_tree_overflow_drops_children_like_the_oracle():
    cfg = _config()
    sc = _comb_scene(cfg.width, cfg.height)
    gpu = Renderer(cfg, sc)
    ref = oracle.RefRenderer(cfg, sc)
    for p in range(2):
        gpu.run_pass()
        gpu.synchronize()
        assert ref.run_pass() == 0
        _compare_pass(gpu, ref, f"comb pass {p}")
    assert gpu.last_photon_count > 0
    seen = _panels_seen(ref.read_hitpoints())
    print("panels seen through the comb:", seen, flush=True)
    assert np.array_equal(seen, np.arange(1, 23)), seen
    # the same scene on the oracle's own (median) tree shows every panel
    full = oracle.RefRenderer(cfg, _comb_scene(cfg.width, cfg.height, reference_tree=False))
    assert full.run_pass() == 0
    assert np.array_equal(_panels_seen(full.read_hitpoints()), np.arange(1, DEPTH + 1))
    full.close()
    gpu.close()


def test_comb_tree_through_the_trace_entry_points():
    """phx_trace_rays and phx_trace_rays_wavefront on the caller's comb nodes."""
    cfg = _config()
    sc = _comb_scene(cfg.width, cfg.height)
    ref = oracle.RefRenderer(cfg, sc)
    assert ref.run_pass() == 0
    ds = DeviceScene(sc, ref)
    nodes = ref.bvh_nodes()
    assert nodes.shape[0] == 2 * DEPTH + 1
    # the pair tree is refused for a tree whose search can overflow
    import ctypes as C
    acc = C.c_uint32(7)
    P.check(P.lib().phx_accel_attach(dptr(ds.nodes), dptr(ds.tv), nodes.ctypes.data, nodes.shape[0],
                                     sc.tri_verts.ctypes.data, ds.n, 0, C.byref(acc)), "attach")
    assert acc.value == 0
    seed = oracle.lib().ref_random_u32(np.array([oracle.lib().ref_pcg_seed(cfg.seed)], np.uint32).ctypes.data)
    n_px = cfg.width * cfg.height
    hp = [empty_dev(16 * n_px) for _ in range(6)]
    L = P.lib()
    P.check(L.phx_trace_rays(sc.camera.ctypes.data, cfg.width, cfg.height, seed, cfg.direct_samples, dptr(ds.tv),
                             dptr(ds.tn), dptr(ds.tt), dptr(ds.tuv), dptr(ds.nodes), dptr(ds.tmat), ds.n,
                             dptr(ds.mats), dptr(ds.uvt), dptr(ds.tex_meta), dptr(ds.atlas), dptr(ds.lights),
                             len(sc.lights), dptr(ds.lum), ds.total_lum, ds.center.ctypes.data, ds.radius,
                             dptr(ds.etv), dptr(ds.etn), dptr(ds.etuv), dptr(ds.etmat), None, 0, 0, 0,
                             *(dptr(b) for b in hp), None), "trace")
    got = np.stack([from_dev(b, np.float32, (n_px, 4)) for b in hp], axis=1)
    want = ref.read_hitpoints()
    assert np.array_equal(_bits(got), _bits(want))
    _check_wavefront(ds, sc, cfg, seed, want, any_hit=0)


def test_thin_lens_camera_passes_bit_exact():
    W, H = 64, 48
    sc = scenes.caustic_scene(W, H, level=3)
    cam = scenes.make_camera((0.5, 0.5, 2.2), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0), 45.0, W / H, defocus_angle=1.5,
                             focus_distance=2.0)
    sc = dataclasses.replace(sc, camera=cam)
    cfg = PassConfig(width=W, height=H, photons_per_round=1 << 14, max_photons_per_batch=1 << 14,
                     max_photon_bounces=12, grid_res=60, seed=9, initial_radius=0.012)
    gpu = Renderer(cfg, sc)
    ref = oracle.RefRenderer(cfg, sc)
    for p in range(2):
        gpu.run_pass()
        gpu.synchronize()
        assert ref.run_pass() == 0
        _compare_pass(gpu, ref, f"thin lens pass {p}")
    # the lens is in use: the camera rays of one pixel start at different points
    assert not np.array_equal(ref.camera_ray(1234, 100)[0], ref.camera_ray(4321, 100)[0])
    gpu.close()
