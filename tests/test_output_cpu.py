"""CPU checks of the output and metric helpers (output.cpp:39-80 tone map,
PNG round trip, bench/metrics.py SSIM/MSE semantics, relative RMSE)."""
import zlib

import numpy as np
import pytest

from phosphor_amd import output


def test_tone_map_values():
    # ACES fit at 0 is 0; identity path is x^(1/2.2) clamped; large values saturate to 1
    assert output.tone_map(np.float32(0.0)) == 0.0
    x = np.float32(0.5)
    assert output.tone_map(x, tonemap=False) == pytest.approx(0.5 ** (1 / 2.2), rel=1e-6)
    a, b, c, d, e = output.ACES
    aces = (1.0 * (a + b)) / (1.0 * (c + d) + e)
    assert output.tone_map(np.float32(1.0)) == pytest.approx(aces ** (1 / 2.2), rel=1e-6)
    assert output.tone_map(np.float32(1e6)) == 1.0
    assert output.tone_map(np.float32(2.0), exposure=0.25, tonemap=False) == pytest.approx(0.5 ** (1 / 2.2), rel=1e-6)
    v = output.tone_map(np.linspace(0, 10, 1000, dtype=np.float32))
    assert np.all(np.diff(v) >= 0)  # monotone
    ldr = output.to_ldr(np.full((2, 2, 3), 1e6, np.float32))
    assert ldr.dtype == np.uint8 and np.all(ldr == 255)


def test_png_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, size=(7, 11, 3), dtype=np.uint8)
    p = output.write_png(tmp_path / "a.png", img, comment="rounds=16 photons=262144")
    back, text = output.read_png(p)
    assert np.array_equal(back, img)
    assert text["Comment"] == "rounds=16 photons=262144"
    raw = p.read_bytes()
    assert raw[:8] == b"\x89PNG\r\n\x1a\n"
    # every chunk's CRC is valid
    pos = 8
    while pos < len(raw):
        n = int.from_bytes(raw[pos:pos + 4], "big")
        tag, body = raw[pos + 4:pos + 8], raw[pos + 8:pos + 8 + n]
        assert int.from_bytes(raw[pos + 8 + n:pos + 12 + n], "big") == zlib.crc32(tag + body) & 0xFFFFFFFF
        pos += 12 + n


def test_metrics():
    rng = np.random.default_rng(4)
    ref = rng.random((32, 40, 3))
    assert output.ssim(ref, ref) == pytest.approx(1.0)
    noisy = np.clip(ref + rng.normal(0, 0.1, ref.shape), 0, 1)
    assert output.ssim(noisy, ref) < 0.99
    assert output.mse(ref + 0.1, ref) == pytest.approx(0.01)
    assert output.relative_rmse(2 * ref, ref) == pytest.approx(1.0)
    # constant images: SSIM is the luminance term (2·μa·μb + C1)/(μa² + μb² + C1)
    a, b = np.full((16, 16, 3), 0.2), np.full((16, 16, 3), 0.4)
    c1 = 0.01 ** 2
    assert output.ssim(a, b) == pytest.approx((2 * 0.2 * 0.4 + c1) / (0.04 + 0.16 + c1))


def test_ssim_matches_direct_window_sum():
    """The separable valid-region filter equals the SSIM definition evaluated
    window by window with the 2-D Gaussian weights (brute force)."""
    rng = np.random.default_rng(0)
    a, b = rng.random((14, 15)), rng.random((14, 15))
    w = np.exp(-0.5 * ((np.arange(11) - 5) / 1.5) ** 2)
    W = np.outer(w, w) / w.sum() ** 2
    vals = []
    for i in range(a.shape[0] - 10):
        for j in range(a.shape[1] - 10):
            pa, pb = a[i:i + 11, j:j + 11], b[i:i + 11, j:j + 11]
            ma, mb = (W * pa).sum(), (W * pb).sum()
            va, vb = (W * pa * pa).sum() - ma * ma, (W * pb * pb).sum() - mb * mb
            c = (W * pa * pb).sum() - ma * mb
            vals.append(((2 * ma * mb + 1e-4) * (2 * c + 9e-4)) / ((ma * ma + mb * mb + 1e-4) * (va + vb + 9e-4)))
    assert output.ssim_channel(a, b) == pytest.approx(np.mean(vals), rel=1e-12)
