"""GPU parity of the photon-map kernels through the C ABI (hash, scan, scatter,
stable sort, gather), against the oracle and closed forms at full sizes."""
import ctypes as C

import numpy as np
import pytest

import phosphor_amd as P
from phosphor_amd.abi import HASH_DTYPE

import oracle

from tests.gpu_util import dptr, empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu


def _hash_info(grid_res, lo=(-0.001, -0.001, -0.001), hi=(1.001, 1.001, 1.001)):
    node = np.zeros((1, 8), np.float32)
    node[0, :3] = lo
    node[0, 4:7] = hi
    d = np.zeros(1, HASH_DTYPE)
    P.check(P.lib().phx_build_photon_hash_info(node.ctypes.data, grid_res, d.ctypes.data))
    ref = np.zeros(1, HASH_DTYPE)
    oracle.lib().ref_build_photon_hash_info(node.ctypes.data, grid_res, ref.ctypes.data)
    assert d.tobytes() == ref.tobytes()
    return d


@pytest.mark.parametrize("n", [0, 1, 4095, 4096, 4097, 262145, 658504, 16581376])
def test_scan_cells(n):
    print(f"scan n={n}", flush=True)
    rng = np.random.default_rng(n)
    counts = rng.integers(0, 40, size=n, dtype=np.uint32)
    if n > 10:
        counts[rng.integers(0, n, size=5)] = 100000  # skewed cells
    d_counts = to_dev(counts)
    d_start = empty_dev(4 * n, 0xAB)
    d_end = empty_dev(4 * n, 0xCD)
    ws_bytes = P.lib().phx_scan_cells_workspace_size(n)
    ws = empty_dev(ws_bytes, 0x5A)
    for rep in range(2):  # the workspace is reused without clearing by the caller
        P.check(P.lib().phx_scan_cells(dptr(d_counts), n, dptr(d_start), dptr(d_end), dptr(ws), ws_bytes, None))
        start = from_dev(d_start, np.uint32, (n,))
        end = from_dev(d_end, np.uint32, (n,))
        expect = np.zeros(n, np.uint32)
        if n:
            expect[1:] = np.cumsum(counts.astype(np.uint64))[:-1].astype(np.uint32)
        assert np.array_equal(start, expect), rep
        assert np.array_equal(end, expect), rep


@pytest.mark.parametrize("grid_res,n", [(43, 50000), (64, 1 << 20), (255, 1 << 22)])
def test_hash_sort_matches_oracle(grid_res, n):
    """hash → scan → stable sort reproduces the oracle's canonical photon map;
    the atomic scatter yields the same per-cell multisets."""
    print(f"hash/sort g={grid_res} n={n}", flush=True)
    rng = np.random.default_rng(grid_res)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-0.05, 1.05, size=(n, 3)).astype(np.float32)  # some outside → key 0
    pos[: n // 10, :3] = np.float32(0.5) + rng.normal(0, 0.002, size=(n // 10, 3)).astype(np.float32)  # hot cells
    power = rng.integers(0, 2**32, size=n, dtype=np.uint32)
    dirs = rng.integers(0, 2**32, size=n, dtype=np.uint32)
    d = _hash_info(grid_res)
    nb = grid_res ** 3 + 1
    # oracle
    h_ref = np.zeros(n, np.uint32)
    c_ref = np.zeros(nb, np.uint32)
    oracle.lib().ref_hash_photons(pos.ctypes.data, n, d.ctypes.data, h_ref.ctypes.data, c_ref.ctypes.data)
    s_ref = np.zeros(nb, np.uint32)
    e_ref = np.zeros(nb, np.uint32)
    oracle.lib().ref_scan_cells(c_ref.ctypes.data, nb, s_ref.ctypes.data, e_ref.ctypes.data)
    sp, spw, sdr = np.zeros_like(pos), np.zeros_like(power), np.zeros_like(dirs)
    oracle.lib().ref_sort_photons(pos.ctypes.data, power.ctypes.data, dirs.ctypes.data, h_ref.ctypes.data,
                                  e_ref.ctypes.data, n, sp.ctypes.data, spw.ctypes.data, sdr.ctypes.data)
    # GPU
    L = P.lib()
    dpos, dpw, ddr = to_dev(pos), to_dev(power), to_dev(dirs)
    dh, dc = empty_dev(4 * n), empty_dev(4 * nb, 0)
    P.check(L.phx_hash_photons(dptr(dpos), n, d.ctypes.data, dptr(dh), dptr(dc), None))
    assert np.array_equal(from_dev(dh, np.uint32, (n,)), h_ref)
    assert np.array_equal(from_dev(dc, np.uint32, (nb,)), c_ref)
    ds, de = empty_dev(4 * nb), empty_dev(4 * nb)
    sws = L.phx_scan_cells_workspace_size(nb)
    dsws = empty_dev(sws)
    P.check(L.phx_scan_cells(dptr(dc), nb, dptr(ds), dptr(de), dptr(dsws), sws, None))
    ows = L.phx_sort_photons_workspace_size(n, nb)
    dows = empty_dev(ows)
    dop, dopw, dodr = empty_dev(16 * n), empty_dev(4 * n), empty_dev(4 * n)
    P.check(L.phx_sort_photons(dptr(dpos), dptr(dpw), dptr(ddr), dptr(dh), dptr(ds), dptr(de), n, nb, dptr(dop),
                               dptr(dopw), dptr(dodr), dptr(dows), ows, None))
    assert np.array_equal(from_dev(ds, np.uint32, (nb,)), s_ref)
    assert np.array_equal(from_dev(de, np.uint32, (nb,)), e_ref)
    assert np.array_equal(from_dev(dop, np.uint32, (n, 4)), sp.view(np.uint32))
    assert np.array_equal(from_dev(dopw, np.uint32, (n,)), spw)
    assert np.array_equal(from_dev(dodr, np.uint32, (n,)), sdr)
    # reference-semantics atomic scatter: same multiset per cell, cell_end = start + count
    P.check(L.phx_scan_cells(dptr(dc), nb, dptr(ds), dptr(de), dptr(dsws), sws, None))
    P.check(L.phx_scatter_photons(dptr(dpos), dptr(dpw), dptr(ddr), dptr(dh), dptr(de), n, dptr(dop), dptr(dopw),
                                  dptr(dodr), None))
    assert np.array_equal(from_dev(de, np.uint32, (nb,)), e_ref)
    got_pw = from_dev(dopw, np.uint32, (n,))
    cell_of = np.repeat(np.arange(nb, dtype=np.int64), c_ref.astype(np.int64))
    key_g = cell_of * (1 << 32) + got_pw
    key_r = cell_of * (1 << 32) + spw
    assert np.array_equal(np.sort(key_g), np.sort(key_r))


def test_error_paths():
    L = P.lib()
    d = np.zeros(1, HASH_DTYPE)
    assert L.phx_hash_photons(None, 10, d.ctypes.data, None, None, None) == P.abi.PHX_ERR_INVALID_ARG
    assert L.phx_scan_cells(None, 10, None, None, None, 0, None) != 0
    assert b"phx_scan_cells" in L.phx_last_error()
