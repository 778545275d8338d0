"""phx_scan_cells_noalloc (workspace in the library's file-scope device array)
against phx_scan_cells (caller workspace) through the C ABI: bit-identical
cell_start and cell_end, and both equal to numpy's exclusive cumsum, on
calls in a row of decreasing size, the first at the entry point's maximum."""
import numpy as np
import pytest

import phosphor_amd as P
from tests.gpu_util import dptr, empty_dev, from_dev, need_gpu, to_dev

pytestmark = pytest.mark.gpu


def _scan_both(counts: np.ndarray):
    L = P.lib()
    n = int(counts.size)
    d_counts = to_dev(counts)
    outs = []
    for noalloc in (False, True):
        d_start = empty_dev(4 * n, fill=0xAB)
        d_end = empty_dev(4 * n, fill=0xCD)
        if noalloc:
            P.check(L.phx_scan_cells_noalloc(dptr(d_counts), n, dptr(d_start), dptr(d_end), None), "scan noalloc")
        else:
            ws_bytes = L.phx_scan_cells_workspace_size(n)
            ws = empty_dev(ws_bytes, fill=0x5A)
            P.check(L.phx_scan_cells(dptr(d_counts), n, dptr(d_start), dptr(d_end), dptr(ws), ws_bytes, None),
                    "scan")
        outs.append((from_dev(d_start, np.uint32, n), from_dev(d_end, np.uint32, n)))
    return outs


def test_scan_noalloc_matches_scan_on_successive_calls():
    need_gpu()
    rng = np.random.default_rng(2024)
    max_cells = 33423360
    for step, n in enumerate((max_cells, 262145, 4097, 1)):
        print(f"[no-alloc scan] call {step}: n_cells={n}", flush=True)
        counts = rng.integers(0, 4, n, dtype=np.uint32)
        if n > 10:
            counts[rng.integers(0, n, 8)] = 100000       # a few hot cells, as in caustic maps
        (s_ref, e_ref), (s_na, e_na) = _scan_both(counts)
        want = np.zeros(n, np.uint64)
        np.cumsum(counts[:-1], dtype=np.uint64, out=want[1:])
        assert np.array_equal(s_ref, want.astype(np.uint32))
        assert np.array_equal(s_na, s_ref), f"cell_start differs at call {step}"
        assert np.array_equal(e_na, e_ref), f"cell_end differs at call {step}"
    L = P.lib()
    assert L.phx_scan_cells_noalloc(1, max_cells + 1, 1, 1, None) == P.abi.PHX_ERR_TOO_LARGE
