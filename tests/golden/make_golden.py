"""Writes tests/golden/kats.npz: known-answer vectors for the pass's integer and
codec primitives, computed here independently of the oracle and of the HIP
library (Python integers and numpy routines), from the definitions the
reference states:

* PCG32 with the RXS-M-XS output (random.h:16-36): seed → state and the first
  outputs, in arbitrary-precision Python integers reduced mod 2^32.
* RGBE (encoding.h:10-39): the exponent from numpy.frexp, channels
  floor(c / 2^e · 256) clamped to 255, E = e + 128; decode c·2^(E-136).
* get_hash (photon_hash.h:21-37): numpy float32 division (IEEE, correctly
  rounded), the dense grid index g²x + gy + z + 1, 0 outside.

The reference ships no golden vectors of its own for this path, so these
fixtures pin the restatement against the definitions, not against reference
runs.  Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "kats.npz"
M32 = (1 << 32) - 1


def pcg_seed(seed: int) -> int:
    s = 0
    s = (s * 747796405 + 2891336453) & M32
    s = (s + seed) & M32
    s = (s * 747796405 + 2891336453) & M32
    return s


def pcg_next(state: int) -> tuple[int, int]:
    old = state
    state = (state * 747796405 + 2891336453) & M32
    word = ((old >> ((old >> 28) + 4)) ^ old) & M32
    return state, (word * 277803737) & M32


def rgbe_encode(p: np.ndarray) -> np.ndarray:
    p = p.astype(np.float32)
    m = np.maximum(np.maximum(p[:, 0], p[:, 1]), p[:, 2])
    out = np.zeros(p.shape[0], np.uint64)
    ok = m > 0
    _, e = np.frexp(m[ok])
    with np.errstate(over="ignore"):  # m ≥ 2^127 gives e = 128 and scale = inf, as ldexp does in C
        scale = np.ldexp(np.float32(1.0), e).astype(np.float32)
    ch = []
    for k in range(3):
        v = (p[ok, k] / scale).astype(np.float32) * np.float32(256.0)
        v = np.where(v > 0, np.floor(v), 0).astype(np.float64)
        ch.append(np.minimum(v, 255).astype(np.uint64))
    E = (e + 128).astype(np.uint64)
    out[ok] = ch[0] | (ch[1] << 8) | (ch[2] << 16) | (E << 24)
    return out.astype(np.uint32)


def rgbe_decode(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64)
    E = (v >> 24).astype(np.int64)
    scale = np.ldexp(np.float32(1.0), E - 136).astype(np.float32)
    out = np.stack([((v >> (8 * k)) & 0xFF).astype(np.float32) * scale for k in range(3)] +
                   [np.zeros(v.shape[0], np.float32)], axis=1)
    out[E == 0] = 0
    return out.astype(np.float32)


def grid_key(pos: np.ndarray, origin: np.ndarray, cell: np.ndarray, g: int) -> np.ndarray:
    f = ((pos[:, :3].astype(np.float32) - origin[:3].astype(np.float32)) / cell[:3].astype(np.float32)).astype(
        np.float32)
    outside = (f < 0).any(axis=1) | (f >= np.float32(g)).any(axis=1)
    c = np.where(outside[:, None], 0, f).astype(np.uint64)
    key = (g * g * c[:, 0] + g * c[:, 1] + c[:, 2] + 1).astype(np.uint64)
    key[outside] = 0
    return key.astype(np.uint32)


def main() -> None:
    rng = np.random.default_rng(20261203)
    seeds = np.array([0, 1, 2, 42, 2137, 0xFFFFFFFF, 0x80000000] + list(rng.integers(0, 2**32, 25)),
                     dtype=np.uint64)
    states = np.array([pcg_seed(int(s)) for s in seeds], np.uint32)
    outs = np.zeros((len(seeds), 16), np.uint32)
    for i, s in enumerate(seeds):
        st = pcg_seed(int(s))
        for k in range(16):
            st, outs[i, k] = pcg_next(st)
    # RGBE: normals, powers of two, values near 2^k boundaries, denormals, zero/negative
    vals = [rng.uniform(0, 10, (200, 4)), rng.uniform(0, 1e-3, (50, 4)), 2.0 ** rng.integers(-140, 100, (50, 4)),
            np.nextafter(2.0 ** rng.integers(-20, 20, (50, 4)).astype(np.float32), np.float32(0)),
            np.array([[0, 0, 0, 0], [-1, -2, -3, 0], [1, 0, 0, 0], [0, 0, 1e-40, 0], [3e38, 1, 0, 0]])]
    rgbe_in = np.concatenate(vals).astype(np.float32)
    rgbe_in[:, 3] = 0
    rgbe_out = rgbe_encode(rgbe_in)
    rgbe_words = np.concatenate([rgbe_out, rng.integers(0, 2**32, 200).astype(np.uint32),
                                 np.array([0, 0x01000000, 0xFF00FFFF, 0x80FFFFFF], np.uint32)])
    rgbe_dec = rgbe_decode(rgbe_words)
    # grid keys, incl. points exactly on cell faces and just outside
    g = 43
    origin = np.array([-0.001, -0.001, -0.001, 0], np.float32)
    cell = np.array([np.float32(1.002) / np.float32(g)] * 3 + [0], np.float32)
    pts = rng.uniform(-0.1, 1.1, (2000, 4)).astype(np.float32)
    faces = (origin[:3] + cell[:3] * rng.integers(0, g + 1, (300, 3)).astype(np.float32)).astype(np.float32)
    pts = np.concatenate([pts, np.concatenate([faces, np.zeros((300, 1), np.float32)], axis=1)])
    pts[:, 3] = 0
    keys = grid_key(pts, origin, cell, g)
    np.savez_compressed(OUT, pcg_seeds=seeds.astype(np.uint32), pcg_states=states, pcg_outputs=outs,
                        rgbe_in=rgbe_in, rgbe_out=rgbe_out, rgbe_words=rgbe_words, rgbe_decoded=rgbe_dec,
                        hash_grid=np.array([g], np.uint32), hash_origin=origin, hash_cell=cell, hash_points=pts,
                        hash_keys=keys)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
