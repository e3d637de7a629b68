"""Independent NumPy float32 implementation of the published noise algorithms,
written from the papers (not from the C oracle), used to cross-check the
oracle bit for bit.

 * Improved Noise: K. Perlin, "Improving Noise", SIGGRAPH 2002 -- quintic
   fade 6t^5 - 15t^4 + 10t^3, 12 edge gradients selected by hash & 15, hash
   chain p[p[p[X]+Y]+Z].  Evaluated vectorised over arrays of points; every
   elementwise float32 operation rounds once (NumPy does not fuse).
 * splitmix64 (Steele, Lea, Flood 2014) for the seeded Fisher-Yates shuffle.
 * The 2-D form with the eight gradients (+-1,+-1), (+-1,0), (0,+-1).
"""
from __future__ import annotations

import numpy as np

f32 = np.float32
MASK64 = (1 << 64) - 1


def splitmix64_stream(seed: int):
    s = seed & MASK64
    while True:
        s = (s + 0x9E3779B97F4A7C15) & MASK64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


def permutation(seed: int) -> np.ndarray:
    p = list(range(256))
    g = splitmix64_stream(seed)
    for i in range(255, 0, -1):
        j = next(g) % (i + 1)
        p[i], p[j] = p[j], p[i]
    return np.array(p + p, np.int64)


def _fade(t):
    return (t * t * t) * (t * (t * f32(6) - f32(15)) + f32(10))


def _lerp(t, a, b):
    return a + t * (b - a)


def _grad3(h, x, y, z):
    h = h & 15
    u = np.where(h < 8, x, y)
    v = np.where(h < 4, y, np.where((h == 12) | (h == 14), x, z))
    return np.where((h & 1) == 0, u, -u) + np.where((h & 2) == 0, v, -v)


def noise3(p: np.ndarray, x, y, z) -> np.ndarray:
    x, y, z = (np.asarray(a, f32) for a in (x, y, z))
    fx, fy, fz = np.floor(x), np.floor(y), np.floor(z)
    X = fx.astype(np.int64) & 255
    Y = fy.astype(np.int64) & 255
    Z = fz.astype(np.int64) & 255
    x, y, z = x - fx, y - fy, z - fz
    u, v, w = _fade(x), _fade(y), _fade(z)
    A = p[X] + Y
    AA, AB = p[A] + Z, p[A + 1] + Z
    B = p[X + 1] + Y
    BA, BB = p[B] + Z, p[B + 1] + Z
    one = f32(1)
    x1, y1, z1 = x - one, y - one, z - one
    return _lerp(w,
                 _lerp(v, _lerp(u, _grad3(p[AA], x, y, z), _grad3(p[BA], x1, y, z)),
                       _lerp(u, _grad3(p[AB], x, y1, z), _grad3(p[BB], x1, y1, z))),
                 _lerp(v, _lerp(u, _grad3(p[AA + 1], x, y, z1), _grad3(p[BA + 1], x1, y, z1)),
                       _lerp(u, _grad3(p[AB + 1], x, y1, z1), _grad3(p[BB + 1], x1, y1, z1))))


def _grad2(h, x, y):
    h = h & 7
    out = np.select([h == 0, h == 1, h == 2, h == 3, h == 4, h == 5, h == 6],
                    [x + y, -x + y, x - y, -x - y, x, -x, y], -y)
    return out.astype(f32)


def noise2(p: np.ndarray, x, y) -> np.ndarray:
    x, y = np.asarray(x, f32), np.asarray(y, f32)
    fx, fy = np.floor(x), np.floor(y)
    X = fx.astype(np.int64) & 255
    Y = fy.astype(np.int64) & 255
    x, y = x - fx, y - fy
    u, v = _fade(x), _fade(y)
    A, B = p[X] + Y, p[X + 1] + Y
    one = f32(1)
    x1, y1 = x - one, y - one
    return _lerp(v, _lerp(u, _grad2(p[A], x, y), _grad2(p[B], x1, y)),
                 _lerp(u, _grad2(p[A + 1], x, y1), _grad2(p[B + 1], x1, y1)))


def fbm3(p, x, y, z, octaves, f0, lac, gain):
    x, y, z = (np.asarray(a, f32) for a in (x, y, z))
    s = np.zeros(np.broadcast(x, y, z).shape, f32)
    amp, f = f32(1), f32(f0)
    for _ in range(octaves):
        s = s + amp * noise3(p, x * f, y * f, z * f)
        f = f32(f * f32(lac))
        amp = f32(amp * f32(gain))
    return s


def fbm2(p, x, y, octaves, f0, lac, gain):
    x, y = np.asarray(x, f32), np.asarray(y, f32)
    s = np.zeros(np.broadcast(x, y).shape, f32)
    amp, f = f32(1), f32(f0)
    for _ in range(octaves):
        s = s + amp * noise2(p, x * f, y * f)
        f = f32(f * f32(lac))
        amp = f32(amp * f32(gain))
    return s
