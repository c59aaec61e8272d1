"""Acquisition geometry on the padded compute grid.

Mirrors the host-side set-up of the reference:

* ``compute_size``       -- source_func::set_compute_size
  (Born/source_func_3d.cpp:37-63): pad the model by nbound = bc_a + 4 on
  each side (plus nboundt + nb + 8 more on axis 3 when ``extra_z``) and
  round each axis up to a multiple of 16;
* ``wavelet_derivative`` -- source_func ctor (Born/source_func_3d.cpp:4-14);
* ``source_points``      -- source_func::get_source_func
  (Born/source_func_3d.cpp:100-133): a 3x3 patch on the plane iz-1 with
  Gaussian weights, trace sample t+4 holds weight * wavelet'(t);
* ``receiver_locs``      -- map_data ctor (Born/map_data_3d.cpp:4-32): cell
  ix + iy n1 + (int(s_z)+4) n1 n2 from fractional grid coordinates;
* ``fractional``         -- data_rtm_3d::get_source_func
  (Born/data_rtm_3d.cpp:52-77): physical position -> grid coordinate with the
  compute-grid origin at -d * nbound (Born/source_func_3d.cpp:91-98).
"""
from __future__ import annotations

import math

import numpy as np


def compute_size(nm, nb, fat=4, extra_z=False, block=16):
    """Padded compute-grid size of a model of size nm = (n1, n2, n3)."""
    nbound = nb + fat
    n = [nm[0] + 2 * nbound, nm[1] + 2 * nbound, nm[2] + 2 * nbound]
    if extra_z:
        n[2] += nb + nb + 2 * fat
    return tuple(int(-(-x // block) * block) for x in n), nbound


def ricker(f0, dt, nt, t0=None):
    """Ricker wavelet of peak frequency f0 (Hz), delayed by t0 (default 1.5/f0)."""
    if t0 is None:
        t0 = 1.5 / f0
    t = np.arange(nt) * dt - t0
    a = (math.pi * f0 * t) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float64)


def wavelet_derivative(w, dt):
    """Forward difference / dt with the last sample zero (Born/source_func_3d.cpp:10-14)."""
    w = np.asarray(w, np.float64)
    out = np.zeros_like(w)
    out[:-1] = (w[1:] - w[:-1]) / dt
    return out


def fractional(pos, d, nbound):
    """Grid coordinate of a physical position: (pos - o)/d with o = -d*nbound."""
    return np.asarray(pos, np.float64) / d + nbound


def source_points(n1, n2, fx, fy, fz, wder, ntr):
    """(locs[9], traces[9][ntr]) of one point source (Born/source_func_3d.cpp:100-133)."""
    ix, iy, iz = int(fx), int(fy), int(fz)
    i3 = iz - 1
    zn = fz - i3
    locs = np.zeros(9, np.int32)
    tr = np.zeros((9, ntr), np.float64)
    nw = min(len(wder), ntr - 4)
    i = 0
    for i2 in range(iy - 1, iy + 2):
        yn = fy - i2
        for i1 in range(ix - 1, ix + 2):
            xn = fx - i1
            locs[i] = i1 + i2 * n1 + i3 * n1 * n2
            tr[i, 4:4 + nw] = math.exp(-zn * zn - xn * xn - yn * yn) * np.asarray(wder[:nw])
            i += 1
    return locs, tr.astype(np.float32)


def receiver_locs(n1, n2, sx, sy, sz):
    """Receiver cells (Born/map_data_3d.cpp:17-29): ix + iy n1 + (int(s_z)+4) n1 n2."""
    ix = np.floor(np.asarray(sx, np.float64)).astype(np.int64)
    iy = np.floor(np.asarray(sy, np.float64)).astype(np.int64)
    iz = np.floor(np.asarray(sz, np.float64)).astype(np.int64) + 4
    return (ix + iy * n1 + iz * n1 * n2).astype(np.int32)


def surface_receivers(n, nbound, rec_depth_cells=0.0, step=1, origin=0):
    """Receivers on every `step`-th (x, y) column of the compute grid at depth
    rec_depth below the model top; returns (locs, nx, ny)."""
    n1, n2, _ = n
    xs = np.arange(origin, n1, step, dtype=np.float64)
    ys = np.arange(origin, n2, step, dtype=np.float64)
    sy, sx = np.meshgrid(ys, xs, indexing="ij")
    sz = np.full(sx.shape, rec_depth_cells + nbound)
    return receiver_locs(n1, n2, sx.ravel(), sy.ravel(), sz.ravel()), len(xs), len(ys)


def cfl_number(vmax, dt, d):
    """v dt / d for the smallest spacing (3-D 8th-order leapfrog limit 0.4529)."""
    return vmax * dt / min(d)
