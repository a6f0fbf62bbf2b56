"""Seeded synthetic inputs for the landscape-evolution timestep.

The reference ships no input rasters (its DEM, climate and bedrock files are
listed as missing blobs), so every workload is generated here:

* fractional-Brownian surfaces by diamond-square on a (2^k+1)^2 lattice,
  displacement sigma_l = sigma_0 * 2^(-H*l) at level l (H = 0.8), cropped to
  the requested side and normalised to [0, 800] m;
* a regional tilt of 0.15 m per 30 m cell (0.5 %) rising with the row index;
* conical pits (depth U(1, 10) m, radius U(2, 8) cells);
* random-walk plateaus of U(20, 200) cells set to the exact height of their
  seed cell;
* a nodata frame (-9999) of width ceil(0.05*side), or an irregular catchment
  outline (star-shaped, radius following a seeded random walk in angle);
* rainfall fields (uniform, lognormal) in mm/yr as the reference expects
  (runoffweight.cu:46 divides by 1000).

Everything is a pure function of the seed (numpy PCG64).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

NODATA = -9999.0
BASE_SEED = 20261202


def _diamond_square(k: int, rng: np.random.Generator, hurst: float = 0.8,
                    sigma0: float = 1.0) -> np.ndarray:
    n = (1 << k) + 1
    z = np.zeros((n, n), dtype=np.float64)
    z[0, 0], z[0, -1], z[-1, 0], z[-1, -1] = rng.normal(0.0, sigma0, 4)
    step = n - 1
    level = 1
    while step > 1:
        half = step // 2
        sig = sigma0 * 2.0 ** (-hurst * level)
        # diamond: square centres
        c = (z[0:n - 1:step, 0:n - 1:step] + z[0:n - 1:step, step::step]
             + z[step::step, 0:n - 1:step] + z[step::step, step::step]) * 0.25
        z[half:n - 1:step, half:n - 1:step] = c + rng.normal(0.0, sig, c.shape)
        # square: edge midpoints, averaging the (up to four) diamond neighbours
        for r0, c0 in ((0, half), (half, 0)):
            rows = np.arange(r0, n, step)
            cols = np.arange(c0, n, step)
            acc = np.zeros((rows.size, cols.size))
            cnt = np.zeros((rows.size, cols.size))
            for dr, dc in ((-half, 0), (half, 0), (0, -half), (0, half)):
                rr = rows + dr
                cc = cols + dc
                vr = (rr >= 0) & (rr < n)
                vc = (cc >= 0) & (cc < n)
                sub = np.zeros((rows.size, cols.size))
                sub[np.ix_(vr, vc)] = z[np.ix_(rr[vr], cc[vc])]
                acc += sub
                cnt += np.outer(vr, vc)
            z[np.ix_(rows, cols)] = acc / cnt + rng.normal(0.0, sig, acc.shape)
        step = half
        level += 1
    return z


def fbm_surface(side: int, rng: np.random.Generator, hurst: float = 0.8) -> np.ndarray:
    k = max(1, math.ceil(math.log2(max(side - 1, 1))))
    z = _diamond_square(k, rng, hurst)[:side, :side].copy()
    lo, hi = z.min(), z.max()
    return (z - lo) / (hi - lo) * 800.0


def add_tilt(z: np.ndarray, per_cell: float = 0.15) -> None:
    z += per_cell * np.arange(z.shape[0], dtype=np.float64)[:, None]


def add_pits(z: np.ndarray, count: int, rng: np.random.Generator, margin: int) -> None:
    h, w = z.shape
    for _ in range(count):
        depth = rng.uniform(1.0, 10.0)
        rad = rng.uniform(2.0, 8.0)
        r0 = int(rng.integers(margin, max(margin + 1, h - margin)))
        c0 = int(rng.integers(margin, max(margin + 1, w - margin)))
        ri = int(math.ceil(rad))
        rs = slice(max(r0 - ri, 0), min(r0 + ri + 1, h))
        cs = slice(max(c0 - ri, 0), min(c0 + ri + 1, w))
        rr, cc = np.mgrid[rs, cs]
        d = np.hypot(rr - r0, cc - c0)
        z[rs, cs] -= depth * np.clip(1.0 - d / rad, 0.0, None)


def add_plateaus(z: np.ndarray, count: int, rng: np.random.Generator, margin: int) -> None:
    h, w = z.shape
    moves = np.array([(0, 1), (1, 0), (0, -1), (-1, 0)])
    for _ in range(count):
        size = int(rng.integers(20, 201))
        r = int(rng.integers(margin, max(margin + 1, h - margin)))
        c = int(rng.integers(margin, max(margin + 1, w - margin)))
        height = z[r, c]
        cells = {(r, c)}
        guard = 0
        while len(cells) < size and guard < 50 * size:
            dr, dc = moves[int(rng.integers(0, 4))]
            r = min(max(r + int(dr), margin), h - margin - 1)
            c = min(max(c + int(dc), margin), w - margin - 1)
            cells.add((r, c))
            guard += 1
        idx = np.array(sorted(cells))
        z[idx[:, 0], idx[:, 1]] = height


def apply_frame(z: np.ndarray, frac: float = 0.05) -> int:
    h, w = z.shape
    f = int(math.ceil(frac * min(h, w)))
    z[:f, :] = NODATA
    z[h - f:, :] = NODATA
    z[:, :f] = NODATA
    z[:, w - f:] = NODATA
    return f


def apply_outline(z: np.ndarray, rng: np.random.Generator, min_margin: int = 2) -> None:
    """Star-shaped catchment: radius r(theta) follows a seeded random walk."""
    h, w = z.shape
    nang = 256
    steps = rng.normal(0.0, 0.04, nang)
    walk = np.cumsum(steps)
    walk -= np.linspace(0.0, walk[-1], nang)            # close the loop
    radius = np.clip(0.38 * (1.0 + walk), 0.18, 0.47)   # in units of min(h, w)
    rr, cc = np.mgrid[0:h, 0:w]
    cy, cx = (h - 1) / 2.0, (w - 1) / 2.0
    dy, dx = rr - cy, cc - cx
    ang = (np.arctan2(dy, dx) + np.pi) / (2 * np.pi) * nang
    i0 = np.floor(ang).astype(int) % nang
    t = ang - np.floor(ang)
    rad = (radius[i0] * (1 - t) + radius[(i0 + 1) % nang] * t) * min(h, w)
    outside = np.hypot(dy, dx) > rad
    z[outside] = NODATA
    z[:min_margin, :] = NODATA
    z[-min_margin:, :] = NODATA
    z[:, :min_margin] = NODATA
    z[:, -min_margin:] = NODATA


@dataclass
class Dem:
    z: np.ndarray          # (H, W) float64, NODATA outside
    rain: np.ndarray       # (H, W) float64 mm/yr, only meaningful on mask cells
    seed: int

    @property
    def shape(self):
        return self.z.shape


def make_dem(side: int, seed: int, pits: int, plateaus: int, outline: bool = False,
             rain: str = "uniform", rain_mm: float = 1000.0, height: int | None = None) -> Dem:
    """One synthetic DEM.  rain: 'uniform' (rain_mm everywhere) or
    'lognormal' (lognormal(0, 0.3) rescaled to mean rain_mm over mask cells)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    h = side if height is None else height
    big = max(side, h)
    z = fbm_surface(big, rng)[:h, :side].copy()
    add_tilt(z)
    margin = int(math.ceil(0.05 * min(side, h))) + 1
    add_pits(z, pits, rng, margin)
    add_plateaus(z, plateaus, rng, margin)
    if outline:
        apply_outline(z, rng)
    else:
        apply_frame(z)
    mask = z != NODATA
    if rain == "uniform":
        r = np.full(z.shape, rain_mm)
    elif rain == "lognormal":
        r = rng.lognormal(0.0, 0.3, z.shape)
        r *= rain_mm / r[mask].mean()
    else:
        raise ValueError(rain)
    return Dem(z=z, rain=r, seed=seed)


def _scaled(n: int, s: int, base: int) -> int:
    """A count given for a base x base DEM, scaled to the area of an s x s one."""
    return max(1, n * s * s // (base * base))


def fbm_512(seed_offset: int = 0, side: int | None = None) -> Dem:
    """512^2 fBm DEM: 40 pits, 10 plateaus, 5 % nodata frame, 1000 mm uniform
    rain.  A smaller side (for tests) scales the pit and plateau counts."""
    s = side or 512
    return make_dem(s, BASE_SEED + 1000 + seed_offset, pits=_scaled(40, s, 512), plateaus=_scaled(10, s, 512))


def fbm_4096(seed_offset: int = 0, side: int | None = None) -> Dem:
    """4096^2 fBm DEM: 400 pits, 100 plateaus, 5 % nodata frame, lognormal
    rain with mean 1000 mm.  A smaller side scales the counts."""
    s = side or 4096
    return make_dem(s, BASE_SEED + 2000 + seed_offset, pits=_scaled(400, s, 4096),
                    plateaus=_scaled(100, s, 4096), rain="lognormal")


def fbm_16384(seed_offset: int = 0, side: int | None = None) -> Dem:
    """16384^2 fBm DEM: 5000 pits, 1000 plateaus, 5 % nodata frame, 1000 mm
    uniform rain.  A smaller side scales the counts."""
    s = side or 16384
    return make_dem(s, BASE_SEED + 5000 + seed_offset, pits=_scaled(5000, s, 16384),
                    plateaus=_scaled(1000, s, 16384))


def catchment_batch(count: int, seed: int, lo: int = 256, hi: int = 2048) -> list[Dem]:
    """Batched uneven catchments: sides log-uniform in [lo, hi], irregular outlines."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sides = np.exp(rng.uniform(math.log(lo), math.log(hi), count)).astype(int)
    out = []
    for i, s in enumerate(sides):
        s = int(max(lo, min(hi, s)))
        pits = max(1, int(40 * s * s / (512 * 512)))
        plats = max(1, int(10 * s * s / (512 * 512)))
        out.append(make_dem(s, seed * 131 + i, pits=pits, plateaus=plats, outline=True))
    return out
