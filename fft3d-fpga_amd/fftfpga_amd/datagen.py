"""Seeded synthetic inputs for the benchmark workloads (numpy, vectorised).

Generator: splitmix64 (S. Vigna).  State k (k = 1, 2, ...) is
seed + k * 0x9E3779B97F4A7C15 mod 2^64, mixed by the published finaliser.
Draw order per point is re then im.

- uniform(n, seed): re, im = 2u - 1 with u = (z >> 40) * 2^-24, exact in FP32.
- gauss(n, seed): two draws (a, b) per point, Box-Muller in FP64:
  u1 = ((a >> 11) + 1) 2^-53, u2 = (b >> 11) 2^-53,
  (re, im) = sqrt(-2 ln u1) (cos 2 pi u2, sin 2 pi u2), rounded to FP32.
- spikes(n, seed, ...): sparse +-amplitude spikes added to a uniform field.
- batch_order(counts, seed): Fisher-Yates shuffle of size classes.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_CHUNK = 1 << 22


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Draws start+1 .. start+count of the splitmix64 stream of `seed`."""
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(npts: int, seed: int) -> np.ndarray:
    """complex64[npts], re/im i.i.d. U[-1,1)."""
    out = np.empty(2 * npts, np.float32)
    for s in range(0, 2 * npts, _CHUNK):
        n = min(_CHUNK, 2 * npts - s)
        z = splitmix64(seed, n, s)
        u = (z >> np.uint64(40)).astype(np.float32) * np.float32(1.0 / 16777216.0)
        out[s:s + n] = np.float32(2.0) * u - np.float32(1.0)
    return out.view(np.complex64)


def gauss(npts: int, seed: int) -> np.ndarray:
    """complex64[npts], re/im i.i.d. N(0,1) (Box-Muller in FP64)."""
    out = np.empty(2 * npts, np.float32)
    step = _CHUNK // 2
    for p in range(0, npts, step):
        n = min(step, npts - p)
        z = splitmix64(seed, 2 * n, 2 * p)
        a, b = z[0::2], z[1::2]
        u1 = ((a >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
        u2 = (b >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        r = np.sqrt(-2.0 * np.log(u1))
        t = 2.0 * np.pi * u2
        out[2 * p:2 * (p + n):2] = (r * np.cos(t)).astype(np.float32)
        out[2 * p + 1:2 * (p + n):2] = (r * np.sin(t)).astype(np.float32)
    return out.view(np.complex64)


def uniform_with_spikes(npts: int, seed: int, fraction: float = 1e-3,
                        amplitude: float = 1e-3) -> np.ndarray:
    """U[-1,1) field plus floor(fraction*npts) spikes of +-amplitude (re and im
    signs drawn independently) at positions drawn from the same stream after
    the 2*npts uniform draws."""
    x = uniform(npts, seed)
    nsp = int(npts * fraction)
    z = splitmix64(seed, 2 * nsp, 2 * npts)
    pos = (z[0::2] % np.uint64(npts)).astype(np.int64)
    sgn_re = np.where((z[1::2] >> np.uint64(63)) & np.uint64(1), -1.0, 1.0).astype(np.float32)
    sgn_im = np.where((z[1::2] >> np.uint64(62)) & np.uint64(1), -1.0, 1.0).astype(np.float32)
    xv = x.view(np.float32)
    np.add.at(xv, 2 * pos, np.float32(amplitude) * sgn_re)
    np.add.at(xv, 2 * pos + 1, np.float32(amplitude) * sgn_im)
    return x


def batch_order(counts: dict, seed: int) -> list:
    """Size list with counts[size] copies of each size, Fisher-Yates shuffled
    with splitmix64(seed): for i = n-1 .. 1, j = z_i mod (i+1), swap(i, j)."""
    items = []
    for size in sorted(counts):
        items += [int(size)] * int(counts[size])
    n = len(items)
    z = splitmix64(seed, max(n - 1, 0))
    for t, i in enumerate(range(n - 1, 0, -1)):
        j = int(z[t] % np.uint64(i + 1))
        items[i], items[j] = items[j], items[i]
    return items


def member_seed(seed: int, index: int) -> int:
    """Seed of the index-th member of a batched workload."""
    return ((seed & 0xFFFFFFFF) << 32) | (index & 0xFFFFFFFF)
