"""Pins of the CPU oracle's primitives and integer stages against independent
computations: committed golden vectors (tests/golden/make_golden.py), numpy
library routines, closed forms and brute force."""
from pathlib import Path

import numpy as np
import pytest

import oracle
from phosphor_amd.abi import HASH_DTYPE

GOLDEN = np.load(Path(__file__).resolve().parent / "golden" / "kats.npz")


def L():
    return oracle.lib()


# ---------------- PCG (random.h:16-36) ----------------

def test_pcg_against_golden_vectors():
    for seed, state, outs in zip(GOLDEN["pcg_seeds"], GOLDEN["pcg_states"], GOLDEN["pcg_outputs"]):
        assert L().ref_pcg_seed(int(seed)) == int(state)
        st = np.array([L().ref_pcg_seed(int(seed))], np.uint32)
        got = [L().ref_random_u32(st.ctypes.data) for _ in range(16)]
        assert got == [int(v) for v in outs]


def test_pcg_output_is_a_bijection_of_the_state():
    """RXS-M-XS is invertible (odd multiplier, xorshift by ≥ 4 bits): undo it and
    recover the pre-step state — a wrong shift or multiplier breaks this."""
    inv_mul = pow(277803737, -1, 1 << 32)
    rng = np.random.default_rng(5)
    for s0 in rng.integers(0, 2**32, 2000):
        st = np.array([s0], np.uint32)
        out = L().ref_random_u32(st.ctypes.data)
        word = (out * inv_mul) & 0xFFFFFFFF
        shift = (word >> 28) + 4  # the top 4 bits survive the xorshift unchanged
        old = word
        for _ in range(8):
            old = word ^ (old >> shift)
        assert old == int(s0)
        assert int(st[0]) == (int(s0) * 747796405 + 2891336453) & 0xFFFFFFFF


def test_random_float_uniform_and_edge():
    st = np.array([L().ref_pcg_seed(2137)], np.uint32)
    x = np.array([L().ref_random_float(st.ctypes.data) for _ in range(200000)])
    hist, _ = np.histogram(x, bins=20, range=(0, 1))
    expected = len(x) / 20
    chi2 = ((hist - expected) ** 2 / expected).sum()
    assert chi2 < 45.0  # 19 dof, p ≈ 1e-3
    assert abs(x.mean() - 0.5) < 0.003
    # u32 → f32 rounds to nearest, so the largest outputs give exactly 1.0f (random.h:36)
    assert np.float32(0xFFFFFFFF) / np.float32(4294967296.0) == np.float32(1.0)


# ---------------- codecs (encoding.h:10-74) ----------------

def test_rgbe_against_golden_vectors():
    for p, expect in zip(GOLDEN["rgbe_in"], GOLDEN["rgbe_out"]):
        assert L().ref_encode_rgbe(np.ascontiguousarray(p).ctypes.data) == int(expect), p
    out = np.zeros(4, np.float32)
    for w, expect in zip(GOLDEN["rgbe_words"], GOLDEN["rgbe_decoded"]):
        L().ref_decode_rgbe(int(w), out.ctypes.data)
        assert np.array_equal(out.view(np.uint32), expect.view(np.uint32)), hex(int(w))


def test_rgbe_quantisation_bound():
    rng = np.random.default_rng(1)
    out = np.zeros(4, np.float32)
    for p in rng.uniform(0, 100, (5000, 4)).astype(np.float32):
        p[3] = 0
        L().ref_decode_rgbe(L().ref_encode_rgbe(p.ctypes.data), out.ctypes.data)
        m = p[:3].max()
        _, e = np.frexp(m)
        step = 2.0 ** (e - 8)
        assert np.all(out[:3] <= p[:3]) and np.all(p[:3] - out[:3] < step)


def test_octahedral_round_trip_error():
    rng = np.random.default_rng(2)
    d = rng.normal(size=(20000, 4)).astype(np.float32)
    d[:, 3] = 0
    d[:6, :3] = [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]
    out = np.zeros(4, np.float32)
    worst = 0.0
    for v in d:
        L().ref_decode_oct(L().ref_encode_oct(np.ascontiguousarray(v).ctypes.data), out.ctypes.data)
        u = v[:3].astype(np.float64) / np.linalg.norm(v[:3].astype(np.float64))
        o = out[:3].astype(np.float64)
        assert abs(np.linalg.norm(o) - 1.0) < 1e-6
        worst = max(worst, float(np.arctan2(np.linalg.norm(np.cross(u, o)), np.dot(u, o))))
    # 16+16-bit truncating octahedral quantisation: one grid step (2/65535) per
    # axis, stretched at most ~3x by the octahedral map → below 2e-4 rad
    assert worst < 2e-4


# ---------------- spatial grid (photon_hash.h:21-50) ----------------

def test_get_hash_against_golden_vectors():
    d = np.zeros(1, HASH_DTYPE)
    d["grid_res"] = int(GOLDEN["hash_grid"][0])
    d["origin"] = GOLDEN["hash_origin"]
    d["cell_sizes"] = GOLDEN["hash_cell"]
    for p, k in zip(GOLDEN["hash_points"], GOLDEN["hash_keys"]):
        assert L().ref_get_hash(np.ascontiguousarray(p).ctypes.data, d.ctypes.data) == int(k)


def test_hash_info_closed_form():
    node = np.array([[-1.5, 0.25, 2.0, 0, 2.5, 0.2505, 6.0, 0]], np.float32)
    d = np.zeros(1, HASH_DTYPE)
    L().ref_build_photon_hash_info(node.ctypes.data, 64, d.ctypes.data)
    assert np.array_equal(d["origin"][0, :3], node[0, :3])
    expect = np.maximum((node[0, 4:7] - node[0, :3]) / np.float32(64), np.float32(0.001)).astype(np.float32)
    assert np.array_equal(d["cell_sizes"][0, :3], expect)
    assert d["cell_sizes"][0, 1] == np.float32(0.001)  # MIN_CELL_SIZE clamp on the flat axis
    assert d["grid_res"][0] == 64


def test_scan_and_stable_scatter_brute_force():
    rng = np.random.default_rng(3)
    n, nb = 30000, 2000
    keys = rng.integers(0, nb, n).astype(np.uint32)
    keys[:5000] = 7  # a hot cell
    counts = np.bincount(keys, minlength=nb).astype(np.uint32)
    start = np.zeros(nb, np.uint32)
    end = np.zeros(nb, np.uint32)
    L().ref_scan_cells(counts.ctypes.data, nb, start.ctypes.data, end.ctypes.data)
    assert np.array_equal(start, np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint32))
    pos = np.zeros((n, 4), np.float32)
    pos[:, 0] = np.arange(n)
    pw = np.arange(n, dtype=np.uint32)
    dr = (np.arange(n, dtype=np.uint32) * 3)
    op, opw, odr = np.zeros_like(pos), np.zeros_like(pw), np.zeros_like(dr)
    L().ref_sort_photons(pos.ctypes.data, pw.ctypes.data, dr.ctypes.data, keys.ctypes.data, end.ctypes.data, n,
                         op.ctypes.data, opw.ctypes.data, odr.ctypes.data)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(opw, pw[order]) and np.array_equal(odr, dr[order])
    assert np.array_equal(end, start + counts)


# ---------------- elementary functions (phx_math.h) ----------------

def _ulp32(x):
    x = np.abs(np.asarray(x, np.float32))
    return np.spacing(np.maximum(x, np.float32(np.finfo(np.float32).tiny))).astype(np.float64)


@pytest.mark.parametrize("which,fn,lo,hi,bound", [
    (0, np.sin, -100.0, 100.0, 2.0), (1, np.cos, -100.0, 100.0, 2.0), (2, np.exp, -87.0, 88.0, 2.0),
    (3, np.log, 1e-30, 1e30, 2.0), (6, np.arccos, -1.0, 1.0, 3.0)])
def test_math_ulp_bounds(which, fn, lo, hi, bound):
    if which == 3:
        x = np.geomspace(lo, hi, 400000).astype(np.float32)
    else:
        x = np.linspace(lo, hi, 400000).astype(np.float32)
    out = np.zeros_like(x)
    L().ref_math(which, x.ctypes.data, None, out.ctypes.data, len(x))
    ref = fn(x.astype(np.float64))
    err = np.abs(out.astype(np.float64) - ref) / _ulp32(ref)
    keep = np.abs(ref) > 1e-3 if which in (0, 1) else np.ones_like(ref, bool)
    assert err[keep].max() <= bound, (which, err[keep].max())


def test_math_atan2_pow_and_special_values():
    rng = np.random.default_rng(4)
    y = rng.normal(size=200000).astype(np.float32)
    x = rng.normal(size=200000).astype(np.float32)
    out = np.zeros_like(x)
    L().ref_math(5, y.ctypes.data, x.ctypes.data, out.ctypes.data, len(x))
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert (np.abs(out - ref) / _ulp32(ref)).max() <= 3.0
    base = rng.uniform(1e-3, 1.0, 200000).astype(np.float32)
    ex = np.full_like(base, 2.2)
    L().ref_math(4, base.ctypes.data, ex.ctypes.data, out.ctypes.data, len(base))
    ref = np.power(base.astype(np.float64), 2.2)
    assert (np.abs(out - ref) / _ulp32(ref)).max() <= 32.0
    sv = np.array([0.0, -0.0], np.float32)
    o = np.zeros(2, np.float32)
    L().ref_math(0, sv.ctypes.data, None, o.ctypes.data, 2)
    assert o[0] == 0 and o[1] == 0
    L().ref_math(1, sv.ctypes.data, None, o.ctypes.data, 2)
    assert np.all(o == 1.0)
    one = np.array([1.0, 0.0], np.float32)
    L().ref_math(3, one.ctypes.data, None, o.ctypes.data, 2)
    assert o[0] == 0.0 and o[1] == -np.inf
    L().ref_math(2, sv.ctypes.data, None, o.ctypes.data, 2)
    assert np.all(o == 1.0)


def test_photon_map_order_is_cell_then_octant_then_deposit():
    """The oracle's sorted photon map equals its deposits ordered by
    (get_hash cell, half-cell octant, deposit index), computed here in numpy
    float32 with get_hash's quotients (photon_hash.h:21-37): octant bit a is
    set where the quotient's fraction is at least 1/2.  Cell ranges are the
    runs of that order.  Also with the octant order off: (cell, deposit)."""
    import oracle
    from phosphor_amd import PassConfig, scenes

    sc = scenes.caustic_scene(48, 40, level=3)
    cfg = PassConfig(width=48, height=40, photons_per_round=20000, max_photons_per_batch=20000,
                     max_photon_bounces=12, grid_res=23, seed=5, initial_radius=0.05)
    for octants in (True, False):
        ref = oracle.RefRenderer(cfg, sc, octant_order=octants)
        assert ref.run_pass() == 0
        pos, pw, dr = ref.read_deposits()
        d = ref.hash_info()[0]
        f = ((pos[:, :3] - d["origin"][:3]) / d["cell_sizes"][:3]).astype(np.float32)
        g = int(d["grid_res"])
        inside = np.all((f >= 0) & (f < np.float32(g)), axis=1)
        c = np.floor(f).astype(np.int64)
        cell = np.where(inside, g * g * c[:, 0] + g * c[:, 1] + c[:, 2] + 1, 0)
        frac = f - c.astype(np.float32)
        octant = np.where(inside, (frac[:, 0] >= 0.5) * 1 + (frac[:, 1] >= 0.5) * 2 + (frac[:, 2] >= 0.5) * 4, 0)
        order = np.lexsort((np.arange(len(cell)), octant if octants else np.zeros_like(octant), cell))
        spos, spw, sdr, cs, ce = ref.read_photon_map()
        assert np.array_equal(spos.view(np.uint32), pos[order].view(np.uint32))
        assert np.array_equal(spw, pw[order]) and np.array_equal(sdr, dr[order])
        counts = np.bincount(cell, minlength=g ** 3 + 1)
        assert np.array_equal(cs, np.concatenate([[0], np.cumsum(counts)[:-1]])) and np.array_equal(ce - cs, counts)
        if octants:
            assert len(np.unique(octant[inside])) == 8
        ref.close()
