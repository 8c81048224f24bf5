"""Seeded synthetic seed sets for the benchmark and parity workloads.

Every set is drawn from SplitMix64 streams so that it can be regenerated
bit for bit anywhere (numpy only, no platform RNG):
  u64 stream  x_i = mix(seed + (i + 1) * 0x9E3779B97F4A7C15)
  uniform     u = (x >> 11) * 2^-53 in [0, 1)
  normal      Box-Muller on two consecutive uniforms
Draws are seed-major: the uniforms of point i are consumed as x, y, z.

Workloads (sizes as named):
  perturbed_grid_32   32^3 = 32768 seeds, perturbed regular grid in the unit
                      cube, offsets uniform in +-0.1 spacing (stream 0),
                      rectangular box -- the recipe of the reference's own
                      perturbed_grid helper (src/voronoi.rs:840-859)
  uniform_2m          128^3 = 2097152 seeds i.i.d. uniform (stream 2),
                      periodic unit cube
  clustered_4m        4194304 seeds (stream 3): 80% in 64 isotropic Gaussian
                      clusters (centres uniform, sigma log-uniform in
                      [1e-3, 1e-1], wrapped), 20% uniform; periodic
  grid_64 / grid_64_jitter   perfect 64^3 grid at cell centres; the jittered
                      copy moves 1% of the seeds (stream 4) by 1e-12*(2u-1)
                      per coordinate
  batched_512         512 problems, n_p = floor(10^U(3,5)) (stream 5); even
                      p uniform, odd p 8 Gaussian clusters (sigma 0.05,
                      wrapped); periodic iff p % 4 < 2
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """u64 outputs offset .. offset+n-1 of the SplitMix64 stream `seed`."""
    i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


class Stream:
    """Sequential uniform/normal draws from one SplitMix64 stream."""

    def __init__(self, seed: int):
        self.seed = seed
        self.pos = 0

    def uniform(self, n: int) -> np.ndarray:
        x = splitmix64(self.seed, n, self.pos)
        self.pos += n
        return (x >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def normal(self, n: int) -> np.ndarray:
        u = self.uniform(2 * n).reshape(n, 2)
        u1 = np.maximum(u[:, 0], 2.0 ** -53)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u[:, 1])

    def integers(self, n: int, hi: int) -> np.ndarray:
        return np.minimum((self.uniform(n) * hi).astype(np.int64), hi - 1)


def perturbed_grid(count: int, pert: float, seed: int = 0, anchor=(0., 0., 0.),
                   width=(1., 1., 1.)) -> np.ndarray:
    """src/voronoi.rs:840-859 with a seeded stream: point n = (i, j, k),
    i = n / count^2, j = (n % count^2) / count, k = n % count;
    pos = (idx + 0.5 + pert * U(-0.5, 0.5)) * width / count + anchor,
    clamped to the box."""
    anchor = np.asarray(anchor, np.float64)
    width = np.asarray(width, np.float64)
    n = count ** 3
    nn = np.arange(n)
    ijk = np.stack([nn // count ** 2, (nn % count ** 2) // count, nn % count], 1).astype(np.float64)
    u = Stream(seed).uniform(3 * n).reshape(n, 3) - 0.5
    pos = (ijk + 0.5 + pert * u) * width / float(count) + anchor
    return np.clip(pos, anchor, anchor + width)


def perturbed_grid_32() -> np.ndarray:
    return perturbed_grid(32, 0.2, seed=0)


def uniform(n: int, seed: int) -> np.ndarray:
    return Stream(seed).uniform(3 * n).reshape(n, 3)


def uniform_2m() -> np.ndarray:
    return uniform(2097152, 2)


def clustered(n: int, seed: int, nclusters: int = 64, frac: float = 0.8,
              sigma_lo: float = 1e-3, sigma_hi: float = 1e-1) -> np.ndarray:
    s = Stream(seed)
    centres = s.uniform(3 * nclusters).reshape(nclusters, 3)
    sig = 10.0 ** (np.log10(sigma_lo) + (np.log10(sigma_hi) - np.log10(sigma_lo)) *
                   s.uniform(nclusters))
    nc = int(round(frac * n))
    which = s.integers(nc, nclusters)
    pts_c = centres[which] + sig[which, None] * s.normal(3 * nc).reshape(nc, 3)
    pts_c = np.mod(pts_c, 1.0)
    pts_u = s.uniform(3 * (n - nc)).reshape(n - nc, 3)
    out = np.concatenate([pts_c, pts_u], 0)
    # np.mod can return exactly 1.0 for tiny negative inputs
    out[out >= 1.0] = 0.0
    return out


def clustered_4m() -> np.ndarray:
    return clustered(4194304, 3)


def grid(count: int) -> np.ndarray:
    nn = np.arange(count ** 3)
    ijk = np.stack([nn // count ** 2, (nn % count ** 2) // count, nn % count], 1)
    return (ijk.astype(np.float64) + 0.5) / count


def grid_jitter(count: int, frac: float = 0.01, seed: int = 4) -> np.ndarray:
    g = grid(count)
    n = g.shape[0]
    k = int(frac * n)
    s = Stream(seed)
    # k distinct indices: order by a uniform key
    pick = np.argsort(s.uniform(n), kind="stable")[:k]
    g[pick] += 1e-12 * (2.0 * s.uniform(3 * k).reshape(k, 3) - 1.0)
    return g


def batched_problems(nprob: int = 512, seed: int = 5, lo: float = 3.0, hi: float = 5.0):
    """List of (positions, periodic) for the batched workload."""
    s = Stream(seed)
    sizes = np.floor(10.0 ** (lo + (hi - lo) * s.uniform(nprob))).astype(np.int64)
    out = []
    for p in range(nprob):
        n = int(sizes[p])
        if p % 2 == 0:
            x = s.uniform(3 * n).reshape(n, 3)
        else:
            c = s.uniform(24).reshape(8, 3)
            w = s.integers(n, 8)
            x = np.mod(c[w] + 0.05 * s.normal(3 * n).reshape(n, 3), 1.0)
            x[x >= 1.0] = 0.0
        out.append((x, (p % 4) < 2))
    return out


def partial_mask(n: int, frac: float, seed: int) -> np.ndarray:
    return Stream(seed).uniform(n) < frac
