"""Seeded synthetic Bluebild workloads (antenna arrays, visibilities, pixel grids).

The visibility model follows the reference simulator: for a sky of point sources
x_k with intensities F_k, the steering vector of source k is
a_k[i] = exp(-i 2pi/lambda p_i . x_k) and

    S = sum_k F_k a_k a_k^H + sigma^2 I        (python/bipp/statistics.py:110-113)

(the Wishart draw of the reference is replaced by the expected matrix, which keeps
the workload deterministic).  Antennas sit uniformly in a disk in the z=0 plane and
rotate about the x-axis by 2pi * (30 t)/86400 rad for timestep t (Earth rotation in
30 s steps).  Flagged antennas have their rows and columns of S zeroed, which is
how the reference's eigensolver sees them (src/eigensolver.cpp:145-168).

Pixel grids follow python/bipp/frame.py:40-72 (half-pixel offset, n = sqrt(1-l^2-m^2)).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

C_LIGHT = 299792458.0


@dataclasses.dataclass
class Sample:
    """One (timestep, channel) sample."""

    t: int
    channel: int
    wl: float
    xyz: np.ndarray        # (N, 3) float64 antenna positions [m]
    s: np.ndarray          # (N, N) complex128 visibilities (flagged rows/cols are zero)
    active: np.ndarray     # (N,) bool, False for flagged antennas


@dataclasses.dataclass
class Workload:
    name: str
    n_antenna: int
    n_levels: int
    grid: int
    fov_deg: float
    samples: List[Sample]
    lmn: np.ndarray        # (P, 3) float64 pixel direction cosines
    src_xyz: np.ndarray    # (K, 3) source directions
    src_flux: np.ndarray   # (K,)
    seed: int
    notes: str = ""

    @property
    def n_samples(self) -> int:
        return len(self.samples)


def make_lmn_grid(grid_size: int, fov_rad: float) -> np.ndarray:
    """(grid_size^2, 3) lmn pixel grid, row-major over (m, l) as frame.make_grids
    flattens it (python/bipp/frame.py:55-72).  Values are rounded to float32, the
    precision in which the reference hands pixel coordinates to synthesis
    (include/bipp/image_prop.hpp:20), and returned as float64."""
    lim = math.sin(fov_rad / 2)
    offset = lim / (grid_size / 2) * 0.5
    s1 = np.linspace(-lim - offset, lim - offset, grid_size)
    s2 = np.linspace(-lim + offset, lim + offset, grid_size)
    l_grid, m_grid = np.meshgrid(s2, s1)
    n_grid = np.sqrt(1 - l_grid ** 2 - m_grid ** 2)
    lmn = np.stack((l_grid, m_grid, n_grid), axis=0).reshape(3, -1).T
    return lmn.astype(np.float32).astype(np.float64)


def disk_positions(rng: np.random.Generator, n: int, diameter_m: float) -> np.ndarray:
    r = 0.5 * diameter_m * np.sqrt(rng.uniform(size=n))
    th = 2 * np.pi * rng.uniform(size=n)
    return np.stack([r * np.cos(th), r * np.sin(th), np.zeros(n)], axis=1)


def zenith_cap_sources(rng: np.random.Generator, k: int, radius_deg: float):
    cos_min = math.cos(math.radians(radius_deg))
    ct = rng.uniform(cos_min, 1.0, size=k)
    st = np.sqrt(1 - ct ** 2)
    ph = 2 * np.pi * rng.uniform(size=k)
    xyz = np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)
    flux = rng.lognormal(0.0, 1.0, size=k)
    return xyz, flux


def rotate_x(p: np.ndarray, theta: float) -> np.ndarray:
    c, s = math.cos(theta), math.sin(theta)
    rot = np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    return p @ rot.T


def visibilities(xyz: np.ndarray, wl: float, src_xyz: np.ndarray, flux: np.ndarray,
                 sigma2: float) -> np.ndarray:
    a = np.exp((-1j * 2 * np.pi / wl) * (xyz @ src_xyz.T))      # (N, K): a_k[i]
    s = (a * flux) @ a.conj().T
    s += sigma2 * np.eye(xyz.shape[0])
    return s


# name, N, diameter [m], n_t, channels [Hz], n_src, levels, grid, N range (flagging)
_CONFIGS = {
    1: dict(name="cfg1-plain-37ant", n=37, diam=2000.0, n_t=50, freqs=[145e6], n_src=20,
            levels=4, grid=256, n_range=None),
    2: dict(name="cfg2-precision-256ant", n=256, diam=10000.0, n_t=200, freqs=[145e6],
            n_src=200, levels=4, grid=1024, n_range=None),
    3: dict(name="cfg3-batching-2000tx8ch", n=256, diam=10000.0, n_t=2000,
            freqs=list(np.linspace(120e6, 190e6, 8)), n_src=200, levels=4, grid=512,
            n_range=(180, 256)),
    4: dict(name="cfg4-nufft-256ant", n=256, diam=10000.0, n_t=500, freqs=[145e6],
            n_src=200, levels=4, grid=2048, n_range=None),
    5: dict(name="cfg5-large-512ant", n=512, diam=40000.0, n_t=1000,
            freqs=list(np.linspace(50e6, 350e6, 16)), n_src=500, levels=6, grid=4096,
            n_range=(400, 512)),
}


def config_spec(cfg: int) -> dict:
    return dict(_CONFIGS[cfg])


def make_workload(cfg: int, n_timesteps: Optional[int] = None, n_channels: Optional[int] = None,
                  grid: Optional[int] = None, seed: Optional[int] = None,
                  fov_deg: float = 5.0, sigma2: float = 0.1,
                  n_antenna: Optional[int] = None, channels: Optional[List[int]] = None,
                  t_step: int = 1) -> Workload:
    """Build configuration `cfg` (1..5), optionally truncated in timesteps/channels/grid.

    "X km disk" is read as a disk of diameter X.  Seed default: 20261202 + cfg.
    channels: indices into the configuration's channel list (default: the first
    n_channels); t_step: timestep stride (timesteps 0, t_step, 2 t_step, ...).
    """
    spec = _CONFIGS[cfg]
    seed = 20261202 + cfg if seed is None else seed
    rng = np.random.default_rng(seed)
    n = spec["n"] if n_antenna is None else n_antenna
    p0 = disk_positions(rng, n, spec["diam"])
    src_xyz, flux = zenith_cap_sources(rng, spec["n_src"], 5.0)
    if channels is not None:
        freqs = [spec["freqs"][c] for c in channels]
    else:
        freqs = spec["freqs"][: (n_channels or len(spec["freqs"]))]
    n_t = spec["n_t"] if n_timesteps is None else n_timesteps
    g = spec["grid"] if grid is None else grid
    samples = []
    for t in range(0, n_t * t_step, t_step):
        theta = 2 * math.pi * (30.0 * t) / 86400.0
        xyz = rotate_x(p0, theta)
        if spec["n_range"] is not None:
            lo, hi = spec["n_range"]
            hi = min(hi, n)
            lo = min(lo, hi)
            n_act = int(rng.integers(lo, hi + 1))
            active = np.zeros(n, dtype=bool)
            active[rng.choice(n, size=n_act, replace=False)] = True
        else:
            active = np.ones(n, dtype=bool)
        for c, f in enumerate(freqs):
            wl = C_LIGHT / f
            s = visibilities(xyz, wl, src_xyz, flux, sigma2)
            s[~active, :] = 0
            s[:, ~active] = 0
            samples.append(Sample(t=t, channel=c, wl=wl, xyz=xyz, s=s, active=active))
    lmn = make_lmn_grid(g, math.radians(fov_deg))
    return Workload(name=spec["name"], n_antenna=n, n_levels=spec["levels"], grid=g,
                    fov_deg=fov_deg, samples=samples, lmn=lmn, src_xyz=src_xyz,
                    src_flux=flux, seed=seed)


def stack_samples(samples: List[Sample], dtype=np.complex64):
    """Pack samples into contiguous batch arrays in the layout the batched C-ABI takes:
    S as (nS, N, N) with each sample column-major, xyz as (nS, 3, N) (column-major N x 3),
    wl as float64."""
    n_s = len(samples)
    n = samples[0].xyz.shape[0]
    s = np.empty((n_s, n, n), dtype=dtype)
    xyz = np.empty((n_s, 3, n), dtype=np.float64)
    wl = np.empty(n_s, dtype=np.float64)
    for k, smp in enumerate(samples):
        s[k] = smp.s.T          # row-major (n_s, col, row) == column-major per sample
        xyz[k] = smp.xyz.T
        wl[k] = smp.wl
    return s, xyz, wl
