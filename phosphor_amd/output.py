"""Output and image metrics of a render.

``tone_map`` / ``to_ldr`` restate src/renderer/output.cpp:39-80 (exposure,
the Narkowicz ACES fit, gamma 1/2.2, clamp, truncation to 8 bits) on the
linear estimate of ``Renderer.estimate()`` (output.cpp:52-76).  ``write_png``
writes an 8-bit RGB PNG with the standard library (zlib); the metadata that
the reference attaches with exiftool (output.cpp:88-99) goes into a tEXt
"Comment" chunk.  ``mse`` / ``ssim`` use the definitions of bench/metrics.py so
numbers compare (11-tap Gaussian window σ 1.5 over the valid region,
stabilisers (0.01·range)² and (0.03·range)², mean over channels); ``relative_rmse`` is ‖L − L_ref‖₂ / ‖L_ref‖₂ on linear images.
"""
from __future__ import annotations

import struct
import zlib
from pathlib import Path

import numpy as np

ACES = (2.51, 0.03, 2.43, 0.59, 0.14)


def tone_map(x: np.ndarray, exposure: float = 1.0, tonemap: bool = True) -> np.ndarray:
    """output.cpp:40-50 in float32: x·exposure, ACES fit, x^(1/2.2), clamp to [0, 1]."""
    x = np.asarray(x, np.float32) * np.float32(exposure)
    out = x
    if tonemap:
        a, b, c, d, e = (np.float32(v) for v in ACES)
        out = (x * (a * x + b)) / (x * (c * x + d) + e)
    with np.errstate(invalid="ignore"):
        out = np.power(out, np.float32(1.0 / 2.2), dtype=np.float32)
    return np.clip(np.nan_to_num(out, nan=0.0), 0.0, 1.0).astype(np.float32)


def to_ldr(linear: np.ndarray, exposure: float = 1.0, tonemap: bool = True) -> np.ndarray:
    """8-bit image of a linear (H, W, 3) estimate (output.cpp:77-79: value·255 truncated)."""
    return (tone_map(linear, exposure, tonemap) * np.float32(255.0)).astype(np.uint8)


def _chunk(tag: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def write_png(path, rgb: np.ndarray, comment: str | None = None) -> Path:
    """Writes an (H, W, 3) uint8 image as PNG (filter 0 on every row)."""
    rgb = np.ascontiguousarray(rgb, np.uint8)
    if rgb.ndim != 3 or rgb.shape[2] != 3:
        raise ValueError("expected an (H, W, 3) uint8 image")
    h, w, _ = rgb.shape
    raw = b"".join(b"\x00" + rgb[y].tobytes() for y in range(h))
    png = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0))
    if comment:
        png += _chunk(b"tEXt", b"Comment\x00" + comment.encode("latin-1", "replace"))
    png += _chunk(b"IDAT", zlib.compress(raw, 6)) + _chunk(b"IEND", b"")
    path = Path(path)
    path.write_bytes(png)
    return path


def read_png(path) -> tuple[np.ndarray, dict]:
    """Reads what write_png writes (8-bit RGB, filter 0): the image and its text chunks."""
    data = Path(path).read_bytes()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG file")
    pos, idat, text = 8, b"", {}
    w = h = 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        tag = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        if tag == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            if depth != 8 or ctype != 2:
                raise ValueError("only 8-bit RGB is read")
        elif tag == b"IDAT":
            idat += body
        elif tag == b"tEXt":
            k, v = body.split(b"\x00", 1)
            text[k.decode("latin-1")] = v.decode("latin-1")
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    if np.any(raw[:, 0] != 0):
        raise ValueError("only filter type 0 is read")
    return raw[:, 1:].reshape(h, w, 3).copy(), text


def snapshot(renderer, path, exposure: float = 1.0, tonemap: bool = True, comment: str | None = None) -> Path:
    """The renderer's current estimate, tone mapped, as a PNG (ImageWriter::submit)."""
    return write_png(path, to_ldr(renderer.estimate(), exposure, tonemap), comment)


# ---- metrics (the definitions of bench/metrics.py, for comparable numbers) ----
def _window(taps: int = 11, sigma: float = 1.5) -> np.ndarray:
    """Normalised Gaussian weights of the SSIM window (Wang et al. 2004)."""
    t = np.arange(taps, dtype=np.float64)
    w = np.exp(-0.5 * ((t - t.mean()) / sigma) ** 2)
    return w / w.sum()


def _local_mean(img: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Separable Gaussian filter over the valid region (no padding)."""
    rows = np.apply_along_axis(lambda r: np.convolve(r, w[::-1], mode="valid"), 1, img)
    return np.apply_along_axis(lambda c: np.convolve(c, w[::-1], mode="valid"), 0, rows)


def ssim_channel(a: np.ndarray, b: np.ndarray, data_range: float = 1.0) -> float:
    """Mean SSIM of one channel: luminance, contrast and structure terms from
    Gaussian-windowed moments, stabilised by (0.01·range)² and (0.03·range)²."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    k1, k2 = (0.01 * data_range) ** 2, (0.03 * data_range) ** 2
    w = _window()
    mu_a, mu_b = _local_mean(a, w), _local_mean(b, w)
    var_a = _local_mean(a * a, w) - mu_a ** 2
    var_b = _local_mean(b * b, w) - mu_b ** 2
    cov = _local_mean(a * b, w) - mu_a * mu_b
    num = (2.0 * mu_a * mu_b + k1) * (2.0 * cov + k2)
    den = (mu_a ** 2 + mu_b ** 2 + k1) * (var_a + var_b + k2)
    return float(np.mean(num / den))


def ssim(img: np.ndarray, ref: np.ndarray, data_range: float = 1.0) -> float:
    """Mean of the three channels' SSIM."""
    return float(np.mean([ssim_channel(img[..., k], ref[..., k], data_range) for k in range(3)]))


def mse(img: np.ndarray, ref: np.ndarray) -> float:
    return float(np.mean((np.asarray(img, np.float64) - np.asarray(ref, np.float64)) ** 2))


def relative_rmse(img: np.ndarray, ref: np.ndarray) -> float:
    """‖L − L_ref‖₂ / ‖L_ref‖₂ over all pixels and channels."""
    d = np.asarray(img, np.float64) - np.asarray(ref, np.float64)
    return float(np.linalg.norm(d) / max(np.linalg.norm(np.asarray(ref, np.float64)), 1e-300))
