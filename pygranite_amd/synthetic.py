"""Synthetic workloads: analytic wind fields and seeded particle sets (numpy only).

These are the inputs of the benchmark configurations and the parity tests:

* ``double_gyre``: 2D double gyre, A=0.1, eps=0.25, omega=2*pi/10 on [0,2]x[0,1].
* ``abc_flow``: 3D time-modulated ABC flow on a periodic [0, 2*pi)^3 grid.
* ``uniform_seeds``: SplitMix64-driven uniform seeds, reproducible across languages.

All fields are returned as float32 stacks shaped like what a ``DataLoader`` hands out slice by
slice: (S, 2, ny, nx) for 2D and (S, 3, nz, ny, nx) for 3D, with x fastest.
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """First n outputs of SplitMix64 started at `seed` (Steele, Lea, Flood 2014)."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, n: int) -> np.ndarray:
    """n doubles in [0, 1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform_seeds(seed: int, n: int, lo, hi) -> np.ndarray:
    """(n, D) float32 points uniform in the box [lo, hi); coordinates interleaved per point."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    d = lo.size
    u = uniform01(seed, n * d).reshape(n, d)
    return (lo + u * (hi - lo)).astype(np.float32)


# ----------------------------------------------------------------------------------------------
# 2D double gyre
# ----------------------------------------------------------------------------------------------
def double_gyre_velocity(x, y, t, A=0.1, eps=0.25, omega=2 * math.pi / 10):
    a = eps * np.sin(omega * t)
    b = 1.0 - 2.0 * eps * np.sin(omega * t)
    f = a * x * x + b * x
    dfdx = 2.0 * a * x + b
    u = -math.pi * A * np.sin(math.pi * f) * np.cos(math.pi * y)
    v = math.pi * A * np.cos(math.pi * f) * np.sin(math.pi * y) * dfdx
    return u, v


def double_gyre(nx=512, ny=256, nslices=11, dtd=1.0):
    """(S, 2, ny, nx) float32 slices at t = 0, dtd, ...; nodes x_i = i*2/(nx-1), y_j = j/(ny-1).

    Returns (wind, grid_scale) with grid_scale = [2/(nx-1), 1/(ny-1)].
    """
    x = np.arange(nx, dtype=np.float64) * (2.0 / (nx - 1))
    y = np.arange(ny, dtype=np.float64) * (1.0 / (ny - 1))
    X, Y = np.meshgrid(x, y)  # (ny, nx)
    out = np.empty((nslices, 2, ny, nx), np.float32)
    for s in range(nslices):
        u, v = double_gyre_velocity(X, Y, s * dtd)
        out[s, 0] = u
        out[s, 1] = v
    return out, [2.0 / (nx - 1), 1.0 / (ny - 1)]


# ----------------------------------------------------------------------------------------------
# 3D ABC flow
# ----------------------------------------------------------------------------------------------
def abc_amplitudes(t):
    m = 1.0 + 0.1 * math.sin(2 * math.pi * t / 8.0)
    return math.sqrt(3.0) * m, math.sqrt(2.0) * m, 1.0 * m


def abc_velocity(x, y, z, t):
    A, B, C = abc_amplitudes(t)
    u = A * np.sin(z) + C * np.cos(y)
    v = B * np.sin(x) + A * np.cos(z)
    w = C * np.sin(y) + B * np.cos(x)
    return u, v, w


def abc_flow(n=128, nslices=16, dtd=1.0):
    """(S, 3, n, n, n) float32 slices of the ABC flow; nodes i*2*pi/n. grid_scale = 2*pi/n."""
    h = 2 * math.pi / n
    c = np.arange(n, dtype=np.float64) * h
    out = np.empty((nslices, 3, n, n, n), np.float32)
    sz, cz = np.sin(c)[:, None, None], np.cos(c)[:, None, None]
    sy, cy = np.sin(c)[None, :, None], np.cos(c)[None, :, None]
    sx, cx = np.sin(c)[None, None, :], np.cos(c)[None, None, :]
    for s in range(nslices):
        A, B, C = abc_amplitudes(s * dtd)
        out[s, 0] = A * sz + C * cy
        out[s, 1] = B * sx + A * cz
        out[s, 2] = C * sy + B * cx
    return out, [h, h, h]


# ----------------------------------------------------------------------------------------------
# Storage rounding helpers (what FP16 / BF16 slices hold once converted on the device)
# ----------------------------------------------------------------------------------------------
def round_fp16(a: np.ndarray) -> np.ndarray:
    return np.asarray(a, np.float32).astype(np.float16).astype(np.float32)


def round_bf16(a: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even to bfloat16, returned as float32 (NaN kept quiet)."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> np.uint64(16)) & np.uint64(1)
    r = ((u + np.uint64(0x7FFF) + lsb) >> np.uint64(16)) << np.uint64(16)
    r = r.astype(np.uint32)
    nan = np.isnan(a)
    r[nan] = (a.view(np.uint32)[nan] | np.uint32(0x00400000)) & np.uint32(0xFFFF0000)
    return r.view(np.float32)


def make_loader_class(base):
    """Returns a DataLoader subclass (of `base`) that serves an in-memory slice stack.

    Semantics follow the reference's loader contract (source/granite/DataLoader.hpp:26-81):
    windfield() returns the current slice as a list of per-component arrays; step() advances
    and returns False once no further slice exists.
    """

    class SliceLoader(base):
        def __init__(self, wind, uplift=None, constants=None, volumes=None):
            base.__init__(self)
            self._wind = wind
            self._uplift = uplift
            self._constants = constants
            self._volumes = volumes
            self._i = 0
            self.calls = {"step": 0, "windfield": 0}

        def step(self):
            self.calls["step"] += 1
            if self._i + 1 < self._wind.shape[0]:
                self._i += 1
                return True
            return False

        def windfield(self):
            self.calls["windfield"] += 1
            return [np.ascontiguousarray(self._wind[self._i, c])
                    for c in range(self._wind.shape[1])]

        def uplift(self):
            return np.ascontiguousarray(self._uplift[self._i])

        def constants(self):
            return {k: np.ascontiguousarray(v[self._i]) for k, v in (self._constants or {}).items()}

        def additionalVolumes(self):
            return {k: np.ascontiguousarray(v[self._i]) for k, v in (self._volumes or {}).items()}

    return SliceLoader
