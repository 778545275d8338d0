"""CPU checks of the drop-in boundary: the C-ABI library loads without a GPU,
exports every entry point include/phosphor_hip.h declares, the struct layouts
match the reference's, and calls fail cleanly (codes, messages) without a
device.  No kernel is launched here."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import phosphor_amd as P
from phosphor_amd import abi

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "phosphor_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(phx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for must in ("phx_trace_rays", "phx_emit_photons", "phx_hash_photons", "phx_scan_cells",
                 "phx_scatter_photons", "phx_sort_photons", "phx_gather_flux", "phx_renderer_create",
                 "phx_renderer_run_pass"):
        assert must in names


def test_library_exports_every_declared_symbol(phx_lib):
    exported = subprocess.run(["nm", "-D", "--defined-only", str(P.LIB_PATH)], capture_output=True, text=True,
                              check=True).stdout
    syms = {line.split()[-1] for line in exported.splitlines() if line.strip()}
    missing = [n for n in declared_functions() if n not in syms]
    assert not missing, missing
    # and the Python binding knows the signature of each
    assert set(declared_functions()) == set(abi.SIGNATURES), set(declared_functions()) ^ set(abi.SIGNATURES)


def test_struct_layouts_match_reference():
    # sizes of include/common/*.h (camera.h:7-22, material.h:36-59, light.h:14-31, photon_hash.h:11-19)
    assert abi.CAMERA_DTYPE.itemsize == 112 and abi.CAMERA_DTYPE.fields["defocus_angle"][1] == 96
    assert abi.MATERIAL_DTYPE.itemsize == 96 and abi.MATERIAL_DTYPE.fields["metallic"][1] == 48
    assert abi.MATERIAL_DTYPE.fields["diff_index"][1] == 72
    assert abi.LIGHT_DTYPE.itemsize == 128 and abi.LIGHT_DTYPE.fields["kind"][1] == 112
    assert abi.HASH_DTYPE.itemsize == 48 and abi.HASH_DTYPE.fields["grid_res"][1] == 32
    assert abi.UV_BLOCK_DTYPE.itemsize == 192 and abi.TEX_META_DTYPE.itemsize == 16
    # the library's own structs, as static_assert'ed in csrc/abi.cpp
    assert C.sizeof(abi.RenderArgs) == 72 and C.sizeof(abi.KernelInfo) == 52


def test_version_and_error_paths_without_device(phx_lib):
    assert phx_lib.phx_version().startswith(b"phosphor")
    d = np.zeros(1, abi.HASH_DTYPE)
    # argument validation happens before any device work
    assert phx_lib.phx_hash_photons(None, 4, d.ctypes.data, None, None, None) == abi.PHX_ERR_INVALID_ARG
    assert b"null" in phx_lib.phx_last_error()
    d["grid_res"] = 300
    assert phx_lib.phx_hash_photons(None, 0, d.ctypes.data, None, None, None) == abi.PHX_ERR_INVALID_ARG
    assert phx_lib.phx_scan_cells(None, 0, None, None, None, 0, None) == abi.PHX_ERR_WORKSPACE
    assert phx_lib.phx_emit_photons_workspace_size(1 << 20, 16) > (1 << 20) * 16 * 24
    assert phx_lib.phx_sort_photons_workspace_size(1 << 20, 262145) >= 4 * 4 * (1 << 20)


def test_renderer_rejects_bad_arguments(phx_lib):
    from phosphor_amd import PassConfig, scenes

    sc = scenes.cornell_box(8, 8)
    bad = PassConfig(width=8, height=8, photons_per_round=16, grid_res=300)
    h = C.c_void_p()
    args, desc = bad.to_c(), sc.desc()
    rc = phx_lib.phx_renderer_create(C.byref(args), C.byref(desc), C.byref(h))
    assert rc == abi.PHX_ERR_INVALID_ARG and not h.value
    assert b"grid_res" in phx_lib.phx_last_error()


def test_bvh_host_builder_matches_oracle(phx_lib, oracle_lib):
    from phosphor_amd import scenes

    for sc in (scenes.cornell_box(8, 8), scenes.caustic_scene(8, 8, level=3)):
        n = sc.n_triangles
        a = np.zeros((2 * n - 1, 8), np.float32)
        b = np.zeros((2 * n - 1, 8), np.float32)
        assert phx_lib.phx_build_bvh_median(sc.tri_verts.ctypes.data, n, a.ctypes.data) == 2 * n - 1
        assert oracle_lib.ref_build_bvh(sc.tri_verts.ctypes.data, n, b.ctypes.data) == 2 * n - 1
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _sah_tree(phx_lib, sc, max_depth=22, max_leaf=1):
    n = sc.n_triangles
    med = np.zeros((2 * n - 1, 8), np.float32)
    assert phx_lib.phx_build_bvh_median(sc.tri_verts.ctypes.data, n, med.ctypes.data) == 2 * n - 1
    sah = np.zeros((2 * n - 1, 8), np.float32)
    order = np.zeros(n, np.uint32)
    got = phx_lib.phx_build_bvh_sah(sc.tri_verts.ctypes.data, n, max_depth, max_leaf, sah.ctypes.data,
                                    order.ctypes.data)
    return med, sah[:got], order, got


def _leaf_range(w):
    return (int(w) & 0x7FFFFFFF) >> 2, (int(w) & 3) + 1


@pytest.mark.parametrize("which,max_leaf", [("cornell", 1), ("caustic", 1), ("room", 1), ("caustic", 2),
                                            ("room", 4)])
def test_sah_tree_structure(phx_lib, which, max_leaf):
    """The SAH tree orders the triangles (a permutation); its leaves cover
    that order with consecutive ranges of 1 .. max_leaf triangles, each leaf
    box the union of its triangles' expanded boxes in the reference tree;
    child boxes nest in parent boxes, interior hints are 0xFFFFFFF8 | (axis |
    4·swap) and the tree stays within the depth limit.  The pairs hold each
    interior node's two child boxes and references."""
    from phosphor_amd import scenes

    sc = {"cornell": lambda: scenes.cornell_box(8, 8), "caustic": lambda: scenes.caustic_scene(8, 8, level=4),
          "room": lambda: scenes.sphere_room(8, 8, 20_000, np.random.default_rng(5))}[which]()
    n = sc.n_triangles
    med, sah, order, got = _sah_tree(phx_lib, sc, max_leaf=max_leaf)
    assert got > 0 and got % 2 == 1
    assert np.array_equal(np.sort(order), np.arange(n, dtype=np.uint32))
    w_min, w_max = sah[:, 3].view(np.uint32), sah[:, 7].view(np.uint32)
    interior = w_min >= 0xFFFFFFF8
    assert interior.sum() == got // 2
    assert np.all((w_min[interior] & 3) <= 2)
    assert np.all(w_min[~interior] & 0x80000000) and np.all(w_max[~interior] == 0)
    med_leaf = med[:, 3].view(np.uint32) != 0xFFFFFFFF
    med_box = np.zeros((n, 6), np.float32)
    med_box[med[med_leaf, 3].view(np.uint32)] = med[med_leaf][:, [0, 1, 2, 4, 5, 6]]
    covered = np.zeros(n, np.int64)
    for i in np.nonzero(~interior)[0]:
        first, count = _leaf_range(w_min[i])
        assert 1 <= count <= max_leaf and first + count <= n
        covered[first:first + count] += 1
        boxes = med_box[order[first:first + count]]
        assert np.array_equal(sah[i, [0, 1, 2]], boxes[:, :3].min(0))
        assert np.array_equal(sah[i, [4, 5, 6]], boxes[:, 3:].max(0))
    assert np.all(covered == 1)
    depth = np.zeros(got, np.int64)
    for i in np.nonzero(interior)[0]:  # preorder: parents come first
        l, r = i + 1, int(w_max[i])
        assert i < r < got
        for c in (l, r):
            depth[c] = depth[i] + 1
            assert np.all(sah[c, :3] >= sah[i, :3]) and np.all(sah[c, 4:7] <= sah[i, 4:7])
    assert depth[interior].max() <= 22
    # the same tree as pairs: each interior node's two child boxes and references
    n_pairs = got // 2
    pairs = np.zeros((max(1, n_pairs), 16), np.float32)
    root = np.zeros(1, np.uint32)
    assert phx_lib.phx_build_bvh_pairs(sah.ctypes.data, got, pairs.ctypes.data, root.ctypes.data) == n_pairs
    pair_of = np.full(got, -1, np.int64)
    pair_of[np.nonzero(interior)[0]] = np.arange(n_pairs)

    def ref(i):
        return int(pair_of[i]) if interior[i] else int(w_min[i])

    assert root[0] == ref(0)
    leaves = []
    for i in np.nonzero(interior)[0]:
        p = pairs[pair_of[i]]
        l, r = i + 1, int(w_max[i])
        assert np.array_equal(p[[0, 1, 2, 4, 5, 6]], sah[l, [0, 1, 2, 4, 5, 6]])
        assert np.array_equal(p[[8, 9, 10, 12, 13, 14]], sah[r, [0, 1, 2, 4, 5, 6]])
        refs = p[[3, 11]].view(np.uint32)
        assert refs[0] == ref(l) and refs[1] == ref(r)
        for x in refs:
            if x & 0x80000000:
                first, count = _leaf_range(x)
                leaves += list(range(first, first + count))
    assert sorted(leaves) == list(range(n))


def test_sah_tree_depth_limit(phx_lib):
    """A limit no balanced tree meets is refused; a tight one is met, also
    with four triangles per leaf (two levels fewer)."""
    from phosphor_amd import scenes

    sc = scenes.caustic_scene(8, 8, level=3)
    for max_leaf, lv in ((1, 0), (4, 2)):
        k = int(np.ceil(np.log2(sc.n_triangles))) - lv  # a balanced tree's deepest interior node is at depth k - 1
        assert _sah_tree(phx_lib, sc, max_depth=k - 2, max_leaf=max_leaf)[3] == 0
        med, sah, order, got = _sah_tree(phx_lib, sc, max_depth=k - 1, max_leaf=max_leaf)
        assert got > 0
        w_min, w_max = sah[:, 3].view(np.uint32), sah[:, 7].view(np.uint32)
        depth = np.zeros(len(sah), np.int64)
        for i in np.nonzero(w_min >= 0xFFFFFFF8)[0]:
            depth[i + 1] = depth[int(w_max[i])] = depth[i] + 1
        assert depth[w_min >= 0xFFFFFFF8].max() == k - 1


def test_sah_tree_rejects_bad_leaf_size(phx_lib):
    from phosphor_amd import scenes

    sc = scenes.cornell_box(8, 8)
    for bad in (0, 5):
        assert _sah_tree(phx_lib, sc, max_leaf=bad)[3] == 0


def test_no_device_reported_here(phx_lib):
    import os

    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    assert phx_lib.phx_device_count() == 0
    assert phx_lib.phx_kernel_info(None, 0) == -abi.PHX_ERR_NO_DEVICE


def test_batch_rejects_bad_arguments(phx_lib):
    from phosphor_amd import PassConfig, scenes

    sc = scenes.cornell_box(8, 8)
    good, bad = PassConfig(width=8, height=8, photons_per_round=16, grid_res=8), \
        PassConfig(width=8, height=8, photons_per_round=16, max_photon_bounces=0)
    args = (abi.RenderArgs * 2)(good.to_c(), bad.to_c())
    descs = (abi.SceneDesc * 2)(sc.desc(), sc.desc())
    h = C.c_void_p()
    rc = phx_lib.phx_batch_create(C.cast(args, C.c_void_p), C.cast(descs, C.c_void_p), 2, None, C.byref(h))
    assert rc == abi.PHX_ERR_INVALID_ARG and not h.value
    assert b"scene 1" in phx_lib.phx_last_error() and b"max_bounces" in phx_lib.phx_last_error()
    assert phx_lib.phx_batch_create(C.cast(args, C.c_void_p), C.cast(descs, C.c_void_p), 0, None,
                                    C.byref(h)) == abi.PHX_ERR_INVALID_ARG
    opt = abi.BatchOptions(0, 0, 0x80, 0)
    assert phx_lib.phx_batch_create(C.cast(args, C.c_void_p), C.cast(descs, C.c_void_p), 1, C.byref(opt),
                                    C.byref(h)) == abi.PHX_ERR_INVALID_ARG


def test_renderer_flags_checked_before_the_device(phx_lib):
    """PHX_PASS_GRAPH, PHX_REFERENCE_BVH, PHX_PASS_SERIAL and PHX_FUSED_TRACE pass the argument
    check (creation then stops at the missing device); any other bit is
    refused first."""
    import os

    from phosphor_amd import PassConfig, scenes

    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    sc = scenes.cornell_box(8, 8)
    desc = sc.desc()
    h = C.c_void_p()
    for flags, want in ((abi.PHX_PASS_SERIAL, abi.PHX_ERR_NO_DEVICE),
                        (abi.PHX_PASS_GRAPH | abi.PHX_REFERENCE_BVH | abi.PHX_PASS_SERIAL, abi.PHX_ERR_NO_DEVICE),
                        (abi.PHX_FUSED_TRACE, abi.PHX_ERR_NO_DEVICE),
                        (0x10, abi.PHX_ERR_INVALID_ARG)):
        a = PassConfig(width=8, height=8, photons_per_round=16, grid_res=8).to_c()
        a.flags = flags
        assert phx_lib.phx_renderer_create(C.byref(a), C.byref(desc), C.byref(h)) == want, flags
        assert not h.value
    opt = abi.BatchOptions(0, 0, abi.PHX_PASS_SERIAL, 0)
    args = (abi.RenderArgs * 1)(PassConfig(width=8, height=8, photons_per_round=16, grid_res=8).to_c())
    descs = (abi.SceneDesc * 1)(desc)
    assert phx_lib.phx_batch_create(C.cast(args, C.c_void_p), C.cast(descs, C.c_void_p), 1, C.byref(opt),
                                    C.byref(h)) == abi.PHX_ERR_NO_DEVICE


def test_launch_shapes_checked_before_the_device(phx_lib):
    """Block sizes other than 64, 128 or 256 threads and gather tile widths
    that are not powers of two up to 64 are refused before any device work;
    the allowed values pass the check."""
    from phosphor_amd import PassConfig, scenes

    sc = scenes.cornell_box(8, 8)
    desc = sc.desc()
    h = C.c_void_p()
    base = dict(width=8, height=8, photons_per_round=16, grid_res=8)
    for kw in (dict(camera_block=100), dict(shadow_block=32), dict(emit_block=512), dict(gather_tile_width=3),
               dict(gather_tile_width=128)):
        a = PassConfig(**base, **kw).to_c()
        assert phx_lib.phx_renderer_create(C.byref(a), C.byref(desc), C.byref(h)) == abi.PHX_ERR_INVALID_ARG, kw
        assert not h.value
    if phx_lib.phx_device_count() == 0:
        for kw in (dict(camera_block=64, shadow_block=256, emit_block=128, gather_tile_width=64),
                   dict(gather_tile_width=1)):
            a = PassConfig(**base, **kw).to_c()
            assert phx_lib.phx_renderer_create(C.byref(a), C.byref(desc), C.byref(h)) == abi.PHX_ERR_NO_DEVICE, kw


def test_u32_index_overflow_refused(phx_lib):
    """Pixel indices, shadow-queue slots (5 NEE events per pixel, each with
    direct_samples light samples) and deposit slots are u32 on the device:
    configurations beyond them are refused with PHX_ERR_TOO_LARGE before any
    allocation, by the renderer and by the batch (whose trace group sizes the
    shadow queue as well)."""
    from phosphor_amd import PassConfig, scenes

    sc = scenes.cornell_box(8, 8)
    desc = sc.desc()
    h = C.c_void_p()
    cases = (dict(width=3840, height=2160, direct_samples=128),     # 5.3 G shadow slots
             dict(width=70000, height=70000, direct_samples=1),     # 4.9 G pixels
             dict(width=8, height=8, photons_per_round=1 << 28, max_photon_bounces=32))
    for kw in cases:
        cfg = PassConfig(**{"photons_per_round": 16, "grid_res": 8, **kw})
        a = cfg.to_c()
        assert phx_lib.phx_renderer_create(C.byref(a), C.byref(desc), C.byref(h)) == abi.PHX_ERR_TOO_LARGE, kw
        assert not h.value
        args = (abi.RenderArgs * 1)(a)
        descs = (abi.SceneDesc * 1)(desc)
        assert phx_lib.phx_batch_create(C.cast(args, C.c_void_p), C.cast(descs, C.c_void_p), 1, None,
                                        C.byref(h)) == abi.PHX_ERR_TOO_LARGE, kw
    # the largest image the shadow queue allows at 128 samples passes the check
    ok = PassConfig(width=2560, height=1310, direct_samples=128, photons_per_round=16, grid_res=8).to_c()
    want = abi.PHX_ERR_NO_DEVICE if phx_lib.phx_device_count() == 0 else None
    if want is not None:
        assert phx_lib.phx_renderer_create(C.byref(ok), C.byref(desc), C.byref(h)) == want



def test_scan_noalloc_limits_checked_before_the_device(phx_lib):
    """phx_scan_cells_noalloc refuses inputs beyond its 64 KiB scratch array
    (PHX_SCAN_NOALLOC_MAX_CELLS) and null pointers before it touches a device."""
    import os

    assert phx_lib.phx_scan_cells_noalloc(1, 33423360 + 1, 1, 1, None) == abi.PHX_ERR_TOO_LARGE
    assert b"PHX_SCAN_NOALLOC_MAX_CELLS" in phx_lib.phx_last_error()
    assert phx_lib.phx_scan_cells_noalloc(1, 0xFFFFFFFF, 1, 1, None) == abi.PHX_ERR_TOO_LARGE
    assert phx_lib.phx_scan_cells_noalloc(None, 10, None, None, None) == abi.PHX_ERR_INVALID_ARG
    assert phx_lib.phx_scan_cells_workspace_size(33423360) == 64 * 1024
    if not os.path.exists("/dev/kfd"):
        assert phx_lib.phx_scan_cells_noalloc(None, 0, None, None, None) != 0   # no device here


def test_accel_eligibility_checks_each_property(phx_lib):
    """phx_accel_eligible accepts the reference's median tree and refuses a
    tree that breaks any property the pair tree's exactness rests on: a
    shrunk leaf box, an interior box not holding a child, more than 22
    interior levels, coordinates beyond 1024, nodes out of preorder, a
    triangle in two leaves, a node count other than 2n - 1."""
    from phosphor_amd import scenes

    from tests.test_oracle_transport import _comb_tree

    def median(tv):
        n = tv.shape[0]
        t = np.zeros((2 * n - 1, 8), np.float32)
        assert phx_lib.phx_build_bvh_median(tv.ctypes.data, n, t.ctypes.data) == 2 * n - 1
        return t

    def verdict(tree, tv, n_nodes=None):
        tree = np.ascontiguousarray(tree, np.float32)
        tv = np.ascontiguousarray(tv, np.float32)
        got = phx_lib.phx_accel_eligible(tree.ctypes.data, tree.shape[0] if n_nodes is None else n_nodes,
                                         tv.ctypes.data, tv.shape[0])
        return got, phx_lib.phx_last_error().decode()

    sc = scenes.caustic_scene(8, 8, level=3)
    tv = sc.tri_verts
    tree = median(tv)
    assert verdict(tree, tv)[0] == 1
    interior = tree[:, 3].view(np.uint32) == 0xFFFFFFFF
    leaf = int(np.nonzero(~interior)[0][0])
    t = tree.copy()
    t[leaf, 0] += np.float32(1e-4)
    assert verdict(t, tv) == (0, "phx_accel_eligible: a leaf box smaller than its triangle's box grown by 1e-3")
    inner = int(np.nonzero(interior)[0][1])
    t = tree.copy()
    t[inner, 0] += np.float32(0.01)
    assert verdict(t, tv) == (0, "phx_accel_eligible: an interior box does not hold its children's boxes")
    tris, comb = _comb_tree(30)
    assert verdict(comb, tris)[1].startswith("phx_accel_eligible: deeper than 22 interior levels")
    far = (tv * np.float32(2000.0)).astype(np.float32)
    assert verdict(median(far), far) == (0, "phx_accel_eligible: coordinates beyond 1024")
    assert verdict(tree, tv, n_nodes=tree.shape[0] - 1)[0] == 0
    # three triangles; median tree: 0 = (leaf 1, interior 2 = (leaf 3, leaf 4))
    small = tv[:3].copy()
    ok = median(small)
    assert verdict(ok, small)[0] == 1
    leaves = np.nonzero(ok[:, 3].view(np.uint32) != 0xFFFFFFFF)[0]
    # 0 = (interior 1 = (leaf 2, leaf 4), leaf 3): valid children, not preorder
    bad = np.zeros_like(ok)
    for i, r in ((0, 3), (1, 4)):
        bad[i] = ok[0]
        bad[i, 7] = np.uint32(r).view(np.float32)
    for i, src in zip((2, 3, 4), leaves):
        bad[i] = ok[src]
    assert verdict(bad, small) == (0, "phx_accel_eligible: nodes not stored in preorder")
    dup = ok.copy()
    dup[leaves[1], 3] = dup[leaves[0], 3]
    dup[leaves[1], :3] = np.minimum(ok[leaves[0], :3], ok[leaves[1], :3])
    dup[leaves[1], 4:7] = np.maximum(ok[leaves[0], 4:7], ok[leaves[1], 4:7])
    assert verdict(dup, small) == (0, "phx_accel_eligible: a triangle index out of range or in two leaves")
