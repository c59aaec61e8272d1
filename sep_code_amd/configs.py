"""Seeded synthetic workloads for the Born RTM path.

Every generator draws from SplitMix64 (``SplitMix64``), never from a
platform RNG, so a seed fixes the model bit for bit on any host.

Workloads (compute grids include the sponge and the 4-cell zero halo):

* ``rtm128``  -- 128^3, d = 10 m, 24-cell sponge, layered v(z) with 4-6
  interfaces in 1500-3500 m/s, reflectivity of sparse +-0.05 spikes on the
  interfaces, Ricker 20 Hz at the grid centre 2 cells below the model top,
  receivers on all 128 x 128 surface columns, dt = 1 ms, nt = 1000;
* ``rtm512``  -- 512^3, d = 12.5 m, 32-cell sponge, layered 1500-4500 m/s
  plus 3 Gaussian anomalies (sigma 20-60 cells, +-15 %), Ricker 15 Hz,
  512 x 512 receivers, dt = 1 ms, nt = 4000, FP32 checkpoints every 50 steps;
* ``lsrtm256`` -- 256^3, d = 12.5 m, dipping layers (2-10 degrees) plus
  sparse reflectivity, 8 shots on a 2 x 4 surface grid, Ricker 15 Hz,
  nt = 2500, Born data plus 1 % Gaussian noise (sigma = 1 % of the RMS of
  the clean data, so |e| / |d_clean| = 1 %; lsrtm.observed_data).

The geometry follows the reference's conventions (model origin at cell
nbound = sponge + 4 of the compute grid, Born/source_func_3d.cpp:91-98;
source plane iz-1, Born/source_func_3d.cpp:118; receivers 4 cells below the
requested depth, Born/map_data_3d.cpp:18).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import geometry as G


class SplitMix64:
    """SplitMix64 stream (Steele, Lea & Flood 2014)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def integers(self, lo: int, hi: int) -> int:
        return lo + self.next_u64() % (hi - lo)

    def uniform_array(self, n: int) -> np.ndarray:
        """n uniforms in [0, 1), vectorised SplitMix64 continuing this stream."""
        k = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + k * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.state = (self.state + n * 0x9E3779B97F4A7C15) & self.MASK
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))

    def normal_array(self, n: int) -> np.ndarray:
        u1 = self.uniform_array(n)
        u2 = self.uniform_array(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2 * math.pi * u2)


@dataclasses.dataclass
class Shot:
    src_locs: np.ndarray      # (9,) int32
    src_traces: np.ndarray    # (9, ntr) float32
    rec_locs: np.ndarray      # (nrec,) int32
    nrx: int
    nry: int


@dataclasses.dataclass
class Workload:
    name: str
    n: tuple                  # (n1, n2, n3) compute grid
    d: tuple
    dt: float
    nt: int
    jt: int
    jimg: int
    bc_a: float
    bc_b: float
    nbound: int
    vel: np.ndarray           # (n3, n2, n1) m/s on the compute grid
    refl: np.ndarray          # (n3, n2, n1) true reflectivity (compute grid)
    shots: list
    ntr: int                  # receiver/source trace length (nt/jt + 8)
    store: str = "ckpt"
    ckpt: int = 50
    cimg: float = 1e-6
    mute_top: int = 0        # model cells below the model top that are muted

    @property
    def V(self):
        return ((self.vel * self.dt) ** 2).astype(np.float32)

    @property
    def ncell(self):
        return self.n[0] * self.n[1] * self.n[2]


def _layered(rng, n, nbound, k_lo, k_hi, v_lo, v_hi):
    n1, n2, n3 = n
    nif = rng.integers(k_lo, k_hi + 1)
    top, bot = nbound + 6, n3 - nbound - 4
    depths = sorted(rng.integers(top, bot) for _ in range(nif))
    vels = sorted(rng.uniform(v_lo, v_hi) for _ in range(nif + 1))
    vz = np.full(n3, vels[0])
    for k, zk in enumerate(depths):
        vz[zk:] = vels[k + 1]
    return vz, depths


def _spikes_on(rng, n, planes, frac=0.05, amp=0.05):
    n1, n2, n3 = n
    refl = np.zeros((n3, n2, n1), np.float32)
    for zk in planes:
        u = rng.uniform_array(n1 * n2).reshape(n2, n1)
        s = rng.uniform_array(n1 * n2).reshape(n2, n1)
        m = u < frac
        refl[zk][m] = np.where(s[m] < 0.5, -amp, amp)
    return refl


def _point_source_shot(n, nbound, wder, ntr, sx_cells, sy_cells, sz_cells, rec_step=1):
    n1, n2, _ = n
    fx, fy, fz = sx_cells + nbound, sy_cells + nbound, sz_cells + nbound
    slocs, strc = G.source_points(n1, n2, fx, fy, fz, wder, ntr)
    rlocs, nx, ny = G.surface_receivers(n, nbound, 0.0, step=rec_step)
    return Shot(slocs, strc, rlocs, nx, ny)


def rtm128(seed: int = 2024, nt: int = 1000) -> Workload:
    n, d, dt, sponge = (128, 128, 128), (10.0, 10.0, 10.0), 0.001, 24
    nbound = sponge + 4
    rng = SplitMix64(seed)
    vz, depths = _layered(rng, n, nbound, 4, 6, 1500.0, 3500.0)
    vel = np.broadcast_to(vz[:, None, None], (n[2], n[1], n[0])).astype(np.float32).copy()
    refl = _spikes_on(rng, n, depths)
    ntr = nt + 8
    w = G.wavelet_derivative(G.ricker(20.0, dt, nt), dt)
    c = n[0] / 2 - nbound
    shot = _point_source_shot(n, nbound, w, ntr, c, c, 2.0)
    return Workload("rtm128", n, d, dt, nt, 1, 1, float(sponge), 0.0015, nbound, vel, refl,
                    [shot], ntr, store="f32")


def rtm512(seed: int = 4096, nt: int = 4000, n_side: int = 512, rec_step: int = 1,
           store: str = "ckpt", jimg: int = 1) -> Workload:
    """The 512^3 production shot; n_side < 512 gives the same physics on a
    smaller grid (sponge, anomaly sizes and depths scaled with the grid)."""
    f = n_side / 512.0
    n, d, dt, sponge = (n_side,) * 3, (12.5, 12.5, 12.5), 0.001, max(8, round(32 * f))
    nbound = sponge + 4
    rng = SplitMix64(seed)
    vz, depths = _layered(rng, n, nbound, 5, 8, 1500.0, 4500.0)
    vel = np.broadcast_to(vz[:, None, None], (n[2], n[1], n[0])).astype(np.float32).copy()
    z = np.arange(n[2], dtype=np.float32)[:, None, None]
    y = np.arange(n[1], dtype=np.float32)[None, :, None]
    x = np.arange(n[0], dtype=np.float32)[None, None, :]
    for _ in range(3):
        cx, cy = rng.uniform(nbound, n[0] - nbound), rng.uniform(nbound, n[1] - nbound)
        cz = rng.uniform(nbound + 40 * f, n[2] - nbound)
        sig = rng.uniform(20.0, 60.0) * f
        amp = rng.uniform(-0.15, 0.15)
        g = np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / np.float32(2 * sig * sig))
        vel *= (1 + np.float32(amp) * g)
    refl = _spikes_on(rng, n, depths)
    ntr = nt + 8
    w = G.wavelet_derivative(G.ricker(15.0, dt, nt), dt)
    c = n[0] / 2 - nbound
    shot = _point_source_shot(n, nbound, w, ntr, c, c, 2.0, rec_step=rec_step)
    name = "rtm512" if n_side == 512 else f"rtm{n_side}p"
    return Workload(name, n, d, dt, nt, 1, jimg, float(sponge), 0.001, nbound, vel, refl,
                    [shot], ntr, store=store, ckpt=50)


def lsrtm(seed: int = 77, n_side: int = 256, nt: int = 2500, nshots=(2, 4), rec_step=2,
          d=12.5, jimg=4, store="f16", anomalies: int = 0) -> Workload:
    n = (n_side, n_side, n_side)
    dd, dt, sponge = (d, d, d), 0.001, max(8, n_side // 12)
    nbound = sponge + 4
    shallow = max(4, n_side // 16)
    rng = SplitMix64(seed)
    n1, n2, n3 = n
    nif = rng.integers(4, 7)
    top, bot = nbound + shallow + 4, n3 - nbound - 8
    z0s = sorted(rng.uniform(top, bot) for _ in range(nif))
    vels = sorted(rng.uniform(1500.0, 3500.0) for _ in range(nif + 1))
    dips = [math.radians(rng.uniform(2.0, 10.0)) for _ in range(nif)]
    azs = [rng.uniform(0, 2 * math.pi) for _ in range(nif)]
    x = np.arange(n1, dtype=np.float64)[None, :] - n1 / 2
    y = np.arange(n2, dtype=np.float64)[:, None] - n2 / 2
    zz = np.arange(n3, dtype=np.float64)[:, None, None]
    vel = np.full((n3, n2, n1), vels[0], np.float32)
    refl = np.zeros((n3, n2, n1), np.float32)
    for k in range(nif):
        surf = z0s[k] + math.tan(dips[k]) * (math.cos(azs[k]) * x + math.sin(azs[k]) * y)
        below = zz >= surf[None]
        vel[below] = vels[k + 1]
        iz = np.clip(np.round(surf).astype(int), 0, n3 - 1)
        sgn = 1.0 if vels[k + 1] > vels[k] else -1.0
        amp = sgn * (vels[k + 1] - vels[k]) / (vels[k + 1] + vels[k])
        jj, ii = np.meshgrid(np.arange(n2), np.arange(n1), indexing="ij")
        refl[iz, jj, ii] += np.float32(amp)
    # Gaussian velocity anomalies (sigma 5-15 % of the grid, +-15 %)
    for _ in range(anomalies):
        cx, cy = rng.uniform(nbound, n1 - nbound), rng.uniform(nbound, n2 - nbound)
        cz = rng.uniform(top, bot)
        sig = rng.uniform(0.05, 0.15) * n_side
        amp = rng.uniform(-0.15, 0.15)
        g = np.exp(-(((x + n1 / 2 - cx) ** 2)[None] + ((y + n2 / 2 - cy) ** 2)[None] +
                     (zz - cz) ** 2) / (2 * sig * sig))
        vel *= (1 + amp * g).astype(np.float32)
    u = rng.uniform_array(n1 * n2 * n3).reshape(n3, n2, n1)
    s = rng.uniform_array(n1 * n2 * n3).reshape(n3, n2, n1)
    sparse = (u < 0.002).astype(np.float32) * np.where(s < 0.5, -0.05, 0.05).astype(np.float32)
    refl += sparse
    # model is zero in the sponge, the halo and the shallow near field of the
    # sources and receivers (which would otherwise dominate the data)
    refl[:nbound + shallow] = 0
    refl[n3 - nbound:] = 0
    refl[:, :nbound] = 0
    refl[:, n2 - nbound:] = 0
    refl[:, :, :nbound] = 0
    refl[:, :, n1 - nbound:] = 0
    ntr = nt + 8
    w = G.wavelet_derivative(G.ricker(15.0, dt, nt), dt)
    shots = []
    nsx, nsy = nshots
    span1, span2 = n1 - 2 * nbound, n2 - 2 * nbound
    for iy in range(nsy):
        for ix in range(nsx):
            sx = span1 * (ix + 0.5) / nsx
            sy = span2 * (iy + 0.5) / nsy
            shots.append(_point_source_shot(n, nbound, w, ntr, sx, sy, 2.0, rec_step=rec_step))
    wl = Workload(f"lsrtm{n_side}", n, dd, dt, nt, 1, jimg, float(sponge), 0.002, nbound, vel,
                  refl, shots, ntr, store=store)
    wl.mute_top = shallow  # model-grid cells muted below the model top
    return wl


def smoke(seed: int = 5) -> Workload:
    """A 48^3, 150-step shot for the quick end-to-end check."""
    n, d, dt, sponge = (48, 48, 48), (10.0, 10.0, 10.0), 0.001, 8
    nbound = sponge + 4
    rng = SplitMix64(seed)
    vz, depths = _layered(rng, n, nbound, 2, 3, 1500.0, 3000.0)
    vel = np.broadcast_to(vz[:, None, None], (n[2], n[1], n[0])).astype(np.float32).copy()
    refl = _spikes_on(rng, n, depths, frac=0.2)
    nt = 150
    ntr = nt + 8
    w = G.wavelet_derivative(G.ricker(30.0, dt, nt), dt)
    c = n[0] / 2 - nbound
    shot = _point_source_shot(n, nbound, w, ntr, c, c, 2.0)
    return Workload("smoke48", n, d, dt, nt, 1, 1, float(sponge), 0.01, nbound, vel, refl,
                    [shot], ntr, store="f32", cimg=1e-3)


WORKLOADS = {"rtm128": rtm128, "rtm512": rtm512, "lsrtm256": lsrtm, "smoke": smoke}
