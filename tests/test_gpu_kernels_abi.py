"""GPU parity of the kernel-level C-ABI entry points phx_trace_rays,
phx_emit_photons and phx_gather_flux, called with the reference's argument
lists (trace_rays.cl:14-40, emit_photons.cl:12-32, gather_flux.cl:5-22) on
buffers uploaded by the test, against the oracle."""
import numpy as np
import pytest

import phosphor_amd as P
from phosphor_amd import PassConfig, scenes

import oracle

from tests.gpu_util import dptr, empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu


class DeviceScene:
    """The scene's buffers on the device, laid out as RenderingBuffers binds them."""

    def __init__(self, sc, ref):
        self.sc = sc
        self.tv, self.tn, self.tt, self.tuv = (to_dev(a) for a in (sc.tri_verts, sc.tri_normals, sc.tri_tangents,
                                                                   sc.tri_uvs))
        self.tmat = to_dev(sc.tri_mat_index)
        self.nodes = to_dev(ref.bvh_nodes())
        self.etv, self.etn, self.etuv, self.etmat = (to_dev(a) for a in (sc.etri_verts, sc.etri_normals,
                                                                         sc.etri_uvs, sc.etri_mat_index))
        self.mats, self.uvt, self.lights = to_dev(sc.materials), to_dev(sc.uv_transforms), to_dev(sc.lights)
        lum = []
        run = np.float32(0)
        for p in sc.lights["power"]:
            v = p[0] * np.float32(0.2126) + p[1] * np.float32(0.7152) + p[2] * np.float32(0.0722)
            run = np.float32(run + max(v, np.float32(0)))
            lum.append(run)
        self.lum_host = np.array(lum, np.float32)
        self.lum = to_dev(self.lum_host)
        self.total_lum = float(self.lum_host[-1])
        self.tex_meta = to_dev(np.zeros(4, np.uint32))
        self.atlas = to_dev(np.zeros(16, np.uint8))
        c, r = ref.scene_sphere()
        self.center = c.astype(np.float32)
        self.radius = r
        self.n = sc.n_triangles


def _check_wavefront(ds, sc, cfg, seed, want, any_hit):
    L = P.lib()
    n_px = cfg.width * cfg.height
    hp = [empty_dev(16 * n_px, 0x7F) for _ in range(6)]
    wsb = L.phx_trace_rays_wavefront_workspace_size(cfg.width, cfg.height, cfg.direct_samples)
    ws = empty_dev(wsb, 0x33)
    for _ in range(2):  # the workspace is reused as is
        P.check(L.phx_trace_rays_wavefront(sc.camera.ctypes.data, cfg.width, cfg.height, seed, cfg.direct_samples,
                                           dptr(ds.tv), dptr(ds.tn), dptr(ds.tt), dptr(ds.tuv), dptr(ds.nodes),
                                           dptr(ds.tmat), ds.n, dptr(ds.mats), dptr(ds.uvt), dptr(ds.tex_meta),
                                           dptr(ds.atlas), dptr(ds.lights), len(sc.lights), dptr(ds.lum),
                                           ds.total_lum, ds.center.ctypes.data, ds.radius, dptr(ds.etv),
                                           dptr(ds.etn), dptr(ds.etuv), dptr(ds.etmat), None, 0, 0, 0,
                                           *(dptr(b) for b in hp), any_hit, dptr(ws), wsb, None), "wavefront")
        got = np.stack([from_dev(b, np.float32, (n_px, 4)) for b in hp], axis=1)
        bad = np.nonzero((got.view(np.uint32) != want.view(np.uint32)).any(axis=(1, 2)))[0]
        assert bad.size == 0, f"{bad.size} pixels differ, first {bad[:5]}"


def test_wavefront_trace_any_hit_on_opaque_scene():
    sc = scenes.cornell_box(64, 48)
    cfg = PassConfig(width=64, height=48, photons_per_round=512, max_photon_bounces=4, grid_res=20, seed=4)
    ref = oracle.RefRenderer(cfg, sc)
    assert ref.run_pass() == 0
    ds = DeviceScene(sc, ref)
    seed = oracle.lib().ref_random_u32(np.array([oracle.lib().ref_pcg_seed(4)], np.uint32).ctypes.data)
    _check_wavefront(ds, sc, cfg, seed, ref.read_hitpoints(), any_hit=1)


def test_trace_emit_gather_entry_points():
    sc = scenes.caustic_scene(48, 40, level=3)
    cfg = PassConfig(width=48, height=40, photons_per_round=6000, max_photons_per_batch=6000, max_photon_bounces=12,
                     grid_res=60, seed=11, initial_radius=0.015)
    ref = oracle.RefRenderer(cfg, sc)
    assert ref.run_pass() == 0  # seed of pass 1 drawn from pcg_seed(11)
    ds = DeviceScene(sc, ref)
    L = P.lib()
    seed = oracle.lib().ref_random_u32(np.array([oracle.lib().ref_pcg_seed(11)], np.uint32).ctypes.data)
    n_px = cfg.width * cfg.height
    hp = [empty_dev(16 * n_px) for _ in range(6)]
    P.check(L.phx_trace_rays(sc.camera.ctypes.data, cfg.width, cfg.height, seed, cfg.direct_samples, dptr(ds.tv),
                             dptr(ds.tn), dptr(ds.tt), dptr(ds.tuv), dptr(ds.nodes), dptr(ds.tmat), ds.n,
                             dptr(ds.mats), dptr(ds.uvt), dptr(ds.tex_meta), dptr(ds.atlas), dptr(ds.lights),
                             len(sc.lights), dptr(ds.lum), ds.total_lum, ds.center.ctypes.data, ds.radius,
                             dptr(ds.etv), dptr(ds.etn), dptr(ds.etuv), dptr(ds.etmat), None, 0, 0, 0,
                             *(dptr(b) for b in hp), None), "trace")
    got = np.stack([from_dev(b, np.float32, (n_px, 4)) for b in hp], axis=1)
    want = ref.read_hitpoints()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    _check_wavefront(ds, sc, cfg, seed, want, any_hit=0)

    # emission in two batches with a global offset, then a truncated buffer
    for max_photons in (cfg.photons_per_round * cfg.max_photon_bounces, 2500):
        dpos, dpw, ddr = empty_dev(16 * max_photons), empty_dev(4 * max_photons), empty_dev(4 * max_photons)
        counter = to_dev(np.zeros(1, np.uint32))
        for off, cnt in ((0, 3500), (3500, 2500)):
            wsb = L.phx_emit_photons_workspace_size(cnt, cfg.max_photon_bounces)
            ws = empty_dev(wsb, None)
            P.check(L.phx_emit_photons(dptr(ds.lights), len(sc.lights), max_photons, off, cnt,
                                       cfg.max_photon_bounces, seed, dptr(ds.nodes), dptr(ds.tv), dptr(ds.tn),
                                       dptr(ds.tt), dptr(ds.tuv), dptr(ds.tmat), ds.n, dptr(ds.etv), dptr(ds.etn),
                                       dptr(ds.etuv), dptr(ds.etmat), dptr(ds.mats), dptr(ds.uvt),
                                       dptr(ds.tex_meta), dptr(ds.atlas), dptr(ds.lum), ds.total_lum,
                                       ds.center.ctypes.data, ds.radius, dptr(dpos), dptr(dpw), dptr(ddr),
                                       dptr(counter), dptr(ws), wsb, None), "emit")
        count = int(from_dev(counter, np.uint32, (1,))[0])
        assert count == ref.last_photon_count
        n = min(count, max_photons)
        wpos, wpw, wdr = ref.read_deposits()  # canonical order, untruncated buffer
        assert np.array_equal(from_dev(dpos, np.uint32, (n, 4)), wpos[:n].view(np.uint32))
        assert np.array_equal(from_dev(dpw, np.uint32, (n,)), wpw[:n])
        assert np.array_equal(from_dev(ddr, np.uint32, (n,)), wdr[:n])

    # gather on the oracle's hitpoints and photon map, from the zero state
    ppos, ppw, pdr, cs, ce = ref.read_photon_map()
    d = ref.hash_info()
    flux = np.zeros((n_px, 4), np.float32)
    flux[:, 3] = np.float32(cfg.initial_radius) ** 2
    dflux, dw, drad = to_dev(flux), to_dev(np.zeros(n_px, np.float32)), to_dev(np.zeros((n_px, 4), np.float32))
    dhp = [to_dev(np.ascontiguousarray(want[:, k])) for k in range(6)]
    dp, dpw2, ddr2, dcs, dce = to_dev(ppos), to_dev(ppw), to_dev(pdr), to_dev(cs), to_dev(ce)
    P.check(L.phx_gather_flux(*(dptr(b) for b in dhp), n_px, dptr(dp), dptr(dpw2), dptr(ddr2), len(ppw), dptr(dcs),
                              dptr(dce), d.ctypes.data, cfg.sppm_alpha, dptr(dflux), dptr(dw), dptr(drad), None),
            "gather")
    rf, rw, rr = ref.read_state()
    assert np.array_equal(from_dev(dflux, np.uint32, (n_px, 4)), rf.view(np.uint32))
    assert np.array_equal(from_dev(dw, np.uint32, (n_px,)), rw.view(np.uint32))
    assert np.array_equal(from_dev(drad, np.uint32, (n_px, 4)), rr.view(np.uint32))


def test_fast_forms_on_uploaded_buffers():
    """phx_accel_attach on the oracle's median tree: phx_trace_rays,
    phx_trace_rays_wavefront and phx_emit_photons then run on the pair tree
    and still equal the oracle bit for bit; phx_sort_photons_by_octant of the
    oracle's deposits equals its octant-ordered photon map, and
    phx_gather_flux_octants on that map equals its state.  After
    phx_accel_detach the reference tree is searched again."""
    import ctypes as C

    sc = scenes.caustic_scene(48, 40, level=3)
    cfg = PassConfig(width=48, height=40, photons_per_round=6000, max_photons_per_batch=6000, max_photon_bounces=12,
                     grid_res=60, seed=11, initial_radius=0.015)
    ref = oracle.RefRenderer(cfg, sc)
    assert ref.run_pass() == 0
    ds = DeviceScene(sc, ref)
    L = P.lib()
    nodes = ref.bvh_nodes()
    acc = C.c_uint32(7)
    P.check(L.phx_accel_attach(dptr(ds.nodes), dptr(ds.tv), nodes.ctypes.data, nodes.shape[0],
                               sc.tri_verts.ctypes.data, ds.n, 0, C.byref(acc)), "attach")
    assert acc.value == 1
    seed = oracle.lib().ref_random_u32(np.array([oracle.lib().ref_pcg_seed(11)], np.uint32).ctypes.data)
    n_px = cfg.width * cfg.height
    want = ref.read_hitpoints()
    for attached in (True, False):
        hp = [empty_dev(16 * n_px) for _ in range(6)]
        P.check(L.phx_trace_rays(sc.camera.ctypes.data, cfg.width, cfg.height, seed, cfg.direct_samples,
                                 dptr(ds.tv), dptr(ds.tn), dptr(ds.tt), dptr(ds.tuv), dptr(ds.nodes), dptr(ds.tmat),
                                 ds.n, dptr(ds.mats), dptr(ds.uvt), dptr(ds.tex_meta), dptr(ds.atlas),
                                 dptr(ds.lights), len(sc.lights), dptr(ds.lum), ds.total_lum, ds.center.ctypes.data,
                                 ds.radius, dptr(ds.etv), dptr(ds.etn), dptr(ds.etuv), dptr(ds.etmat), None, 0, 0, 0,
                                 *(dptr(b) for b in hp), None), "trace")
        got = np.stack([from_dev(b, np.float32, (n_px, 4)) for b in hp], axis=1)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"attached={attached}"
        _check_wavefront(ds, sc, cfg, seed, want, any_hit=0)
        max_photons = cfg.photons_per_round * cfg.max_photon_bounces
        dpos, dpw, ddr = empty_dev(16 * max_photons), empty_dev(4 * max_photons), empty_dev(4 * max_photons)
        counter = to_dev(np.zeros(1, np.uint32))
        wsb = L.phx_emit_photons_workspace_size(cfg.photons_per_round, cfg.max_photon_bounces)
        ws = empty_dev(wsb, None)
        P.check(L.phx_emit_photons(dptr(ds.lights), len(sc.lights), max_photons, 0, cfg.photons_per_round,
                                   cfg.max_photon_bounces, seed, dptr(ds.nodes), dptr(ds.tv), dptr(ds.tn),
                                   dptr(ds.tt), dptr(ds.tuv), dptr(ds.tmat), ds.n, dptr(ds.etv), dptr(ds.etn),
                                   dptr(ds.etuv), dptr(ds.etmat), dptr(ds.mats), dptr(ds.uvt), dptr(ds.tex_meta),
                                   dptr(ds.atlas), dptr(ds.lum), ds.total_lum, ds.center.ctypes.data, ds.radius,
                                   dptr(dpos), dptr(dpw), dptr(ddr), dptr(counter), dptr(ws), wsb, None), "emit")
        n = int(from_dev(counter, np.uint32, (1,))[0])
        assert n == ref.last_photon_count
        wpos, wpw, wdr = ref.read_deposits()
        assert np.array_equal(from_dev(dpos, np.uint32, (n, 4)), wpos[:n].view(np.uint32))
        assert np.array_equal(from_dev(dpw, np.uint32, (n,)), wpw[:n])
        assert np.array_equal(from_dev(ddr, np.uint32, (n,)), wdr[:n])
        if attached:
            P.check(L.phx_accel_detach(dptr(ds.nodes)), "detach")

    # the octant-ordered map in one call, then the octant gather
    d = ref.hash_info()
    buckets = cfg.grid_res ** 3 + 1
    spos, spw, sdr, sdp = empty_dev(16 * n), empty_dev(4 * n), empty_dev(4 * n), empty_dev(16 * n)
    dcs, dce, doct = empty_dev(4 * buckets), empty_dev(4 * buckets), empty_dev(32 * buckets, 0xFF)
    wsb = L.phx_sort_photons_by_octant_workspace_size(n, buckets)
    ws = empty_dev(wsb, None)
    P.check(L.phx_sort_photons_by_octant(dptr(dpos), dptr(dpw), dptr(ddr), n, d.ctypes.data, buckets, dptr(spos),
                                         dptr(spw), dptr(sdr), dptr(sdp), dptr(dcs), dptr(dce), dptr(doct), dptr(ws),
                                         wsb, None), "sort_by_octant")
    ppos, ppw, pdr, cs, ce = ref.read_photon_map()
    assert np.array_equal(from_dev(spos, np.uint32, (n, 4)), ppos.view(np.uint32))
    assert np.array_equal(from_dev(spw, np.uint32, (n,)), ppw)
    assert np.array_equal(from_dev(sdr, np.uint32, (n,)), pdr)
    assert np.array_equal(from_dev(dcs, np.uint32, (buckets,)), cs)
    assert np.array_equal(from_dev(dce, np.uint32, (buckets,)), ce)
    flux = np.zeros((n_px, 4), np.float32)
    flux[:, 3] = np.float32(cfg.initial_radius) ** 2
    dflux, dw, drad = to_dev(flux), to_dev(np.zeros(n_px, np.float32)), to_dev(np.zeros((n_px, 4), np.float32))
    dhp = [to_dev(np.ascontiguousarray(want[:, k])) for k in range(6)]
    P.check(L.phx_gather_flux_octants(*(dptr(b) for b in dhp), cfg.width, cfg.height, dptr(spos), dptr(sdp), n,
                                      dptr(dcs), dptr(dce), dptr(doct), d.ctypes.data, cfg.sppm_alpha, dptr(dflux),
                                      dptr(dw), dptr(drad), None), "gather_octants")
    rf, rw, rr = ref.read_state()
    assert np.array_equal(from_dev(dflux, np.uint32, (n_px, 4)), rf.view(np.uint32))
    assert np.array_equal(from_dev(dw, np.uint32, (n_px,)), rw.view(np.uint32))
    assert np.array_equal(from_dev(drad, np.uint32, (n_px, 4)), rr.view(np.uint32))


def test_kernel_info_reports_registers_and_occupancy():
    """phx_kernel_info describes the traversal, emission and photon-map
    kernels: VGPRs within the 512-register file, at least one resident wave
    per SIMD, the block sizes the launches use."""
    info = P.kernel_info()
    print(info)
    for name in ("k_trace_camera", "k_trace_camera_rest", "k_trace_shadow", "k_trace_resolve", "k_emit_wf_bounce",
                 "k_emit_wf_finish", "k_emit_trace", "k_emit_shade", "k_hash", "k_radix_scatter", "k_decode_photons",
                 "k_gather_flat"):
        k = info[name]
        assert 0 < k["vgprs"] <= 512 and k["waves_per_simd"] >= 1, (name, k)
    assert info["k_trace_shadow"]["block"] == 128 and info["k_emit_wf_bounce"]["block"] == 128


@pytest.mark.parametrize("fast", [False, True], ids=["reference_kernels", "fast_forms"])
@pytest.mark.parametrize("case", ["caustic", "atrium"])
def test_stage_pass_matches_the_oracle(case, fast):
    """The reference's pass issued through the stage-level ABI (StagePass:
    trace_rays_wavefront, emit_photons in batches, hash, scan, sort, gather on
    one stream, the reference's median tree) equals the oracle's pass with a
    cell-ordered photon map after each of three passes, bit for bit.  With
    the fast forms (the pair tree attached to the median tree, the map sorted
    by octant in one call, the octant gather) it equals the oracle's pass
    with an octant-ordered map, and the trace entry points do run on the
    pair tree."""
    from phosphor_amd.stages import StagePass

    if case == "caustic":
        sc = scenes.caustic_scene(96, 64, level=4)
        cfg = PassConfig(width=96, height=64, photons_per_round=1 << 15, max_photons_per_batch=12000,
                         max_photon_bounces=16, grid_res=87, seed=7, initial_radius=0.01)
    else:
        sc = scenes.atrium_scene(160, 90, target_tris=40_000, env_size=(256, 128))
        cfg = PassConfig(width=160, height=90, photons_per_round=1 << 15, max_photons_per_batch=1 << 15,
                         max_photon_bounces=16, grid_res=64, seed=3)
    print(f"case stage pass {case} fast={fast}: {sc.n_triangles} triangles", flush=True)
    sp = StagePass(cfg, sc, fast=fast)
    assert sp.accelerated == fast
    ref = oracle.RefRenderer(cfg, sc, octant_order=fast)
    for p in range(3):
        sp.run_pass()
        assert ref.run_pass() == 0
        assert sp.last_photon_count == ref.last_photon_count
        for a, b, name in zip(sp.read_photon_map(), ref.read_photon_map(), ("pos", "power", "dir", "cs", "ce")):
            assert np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32)), \
                f"pass {p}: photon map {name}"
        for a, b, name in zip(sp.read_state(), ref.read_state(), ("flux", "weight", "radiance")):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"pass {p}: {name}"
    assert np.all(sp.stage_ms > 0)
    sp.close()
