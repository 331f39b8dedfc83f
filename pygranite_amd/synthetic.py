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
# Atmosphere-like field: Gaussian random field (k^-5/3 horizontal energy spectrum) + zonal jet
# ----------------------------------------------------------------------------------------------
def stretched_levels(nz=64, dz0=20.0, ratio=1.05):
    """Heights z_k = dz0 (ratio^k - 1) / (ratio - 1), k = 0..nz-1 (z_0 = 0)."""
    k = np.arange(nz, dtype=np.float64)
    return (dz0 * (ratio ** k - 1.0) / (ratio - 1.0)).astype(np.float32)


def _grf_amplitude(xp, nz, ny, nx, dx, levels_dz):
    """Spectral amplitude |u_hat| ~ k_h^(-4/3) (E(k) ~ k^-5/3 in 2D), smooth in z."""
    ky = xp.fft.fftfreq(ny, d=dx)
    kx = xp.fft.rfftfreq(nx, d=dx)
    kz = xp.fft.fftfreq(nz, d=1.0)
    kh = xp.sqrt(ky[:, None] ** 2 + kx[None, :] ** 2)
    kmin = 1.0 / (max(nx, ny) * dx)
    amp_h = xp.where(kh > 0, (xp.maximum(kh, kmin) / kmin) ** (-4.0 / 3.0), 0.0)
    amp_z = xp.exp(-(kz / 0.08) ** 2)  # vertical correlation over ~ a dozen levels
    return amp_z[:, None, None] * amp_h[None, :, :]


def _grf(xp, rng_normal, amp, shape, rms):
    nz, ny, nx = shape
    spec = (rng_normal((nz, ny, nx // 2 + 1)) + 1j * rng_normal((nz, ny, nx // 2 + 1))) * amp
    f = xp.fft.irfftn(spec, s=(nz, ny, nx), axes=(0, 1, 2))
    f = f - f.mean()
    return f * (rms / xp.sqrt(xp.mean(f * f)))


def atmosphere(nx=512, ny=512, nz=64, nslices=25, seed=1, dx=3000.0, dz0=20.0, ratio=1.05,
               device=None):
    """Wind slices (S, 3, nz, ny, nx) float32 of an atmosphere-like flow.

    Recipe: for u, v two independent GRFs A, B with horizontal spectrum E(k) ~ k^-5/3 (RMS
    10 m/s, correlated over ~12 levels in z); slice s (hourly) is cos(th) A + sin(th) B with
    th = 2 pi s / 48, which keeps the RMS and the spectrum and decorrelates over a day. A zonal
    jet 20 cos^2(pi (y - yc) / Ly) m/s is added to u. w is built the same way with RMS
    0.05 m/s. Vertical nodes sit on `stretched_levels(nz, dz0, ratio)`. With `device` set (a
    torch device), the field is generated there with torch and returned as a torch tensor;
    otherwise numpy on the host.

    Returns (wind, grid_scale [dx, dx, mean dz], z_levels).
    """
    levels = stretched_levels(nz, dz0, ratio)
    if device is not None:
        import torch
        xp = torch
        g = torch.Generator(device=device)
        g.manual_seed(int(seed))

        def normal(shape):
            return torch.randn(shape, generator=g, device=device, dtype=torch.float64)

        ky = torch.fft.fftfreq(ny, d=dx, device=device, dtype=torch.float64)
        kx = torch.fft.rfftfreq(nx, d=dx, device=device, dtype=torch.float64)
        kz = torch.fft.fftfreq(nz, d=1.0, device=device, dtype=torch.float64)
        kh = torch.sqrt(ky[:, None] ** 2 + kx[None, :] ** 2)
        kmin = 1.0 / (max(nx, ny) * dx)
        amp_h = torch.where(kh > 0, (torch.clamp(kh, min=kmin) / kmin) ** (-4.0 / 3.0),
                            torch.zeros_like(kh))
        amp = torch.exp(-(kz / 0.08) ** 2)[:, None, None] * amp_h[None, :, :]

        def grf(rms):
            spec = torch.complex(normal((nz, ny, nx // 2 + 1)), normal((nz, ny, nx // 2 + 1))) * amp
            f = torch.fft.irfftn(spec, s=(nz, ny, nx), dim=(0, 1, 2))
            f = f - f.mean()
            return (f * (rms / torch.sqrt(torch.mean(f * f)))).to(torch.float32)

        y = torch.arange(ny, device=device, dtype=torch.float32) * dx
        Ly = ny * dx
        jet = (20.0 * torch.cos(math.pi * (y - Ly / 2) / Ly) ** 2)[None, :, None]
        base = [(grf(10.0), grf(10.0)), (grf(10.0), grf(10.0)), (grf(0.05), grf(0.05))]
        out = torch.empty((nslices, 3, nz, ny, nx), device=device, dtype=torch.float32)
        for s in range(nslices):
            th = 2 * math.pi * s / 48.0
            for c in range(3):
                a, b = base[c]
                out[s, c] = math.cos(th) * a + math.sin(th) * b
            out[s, 0] += jet
        return out, [dx, dx, float(levels[-1]) / (nz - 1)], levels

    rng = np.random.default_rng(seed)
    amp = _grf_amplitude(np, nz, ny, nx, dx, None)
    normal = rng.standard_normal
    base = [(_grf(np, normal, amp, (nz, ny, nx), 10.0), _grf(np, normal, amp, (nz, ny, nx), 10.0)),
            (_grf(np, normal, amp, (nz, ny, nx), 10.0), _grf(np, normal, amp, (nz, ny, nx), 10.0)),
            (_grf(np, normal, amp, (nz, ny, nx), 0.05), _grf(np, normal, amp, (nz, ny, nx), 0.05))]
    y = np.arange(ny, dtype=np.float64) * dx
    Ly = ny * dx
    jet = (20.0 * np.cos(math.pi * (y - Ly / 2) / Ly) ** 2)[None, :, None]
    out = np.empty((nslices, 3, nz, ny, nx), np.float32)
    for s in range(nslices):
        th = 2 * math.pi * s / 48.0
        for c in range(3):
            a, b = base[c]
            out[s, c] = math.cos(th) * a + math.sin(th) * b
        out[s, 0] += jet
    return out, [dx, dx, float(levels[-1]) / (nz - 1)], levels


def loguniform_heights(seed, n, lo, hi):
    u = uniform01(seed, n)
    return (lo * np.exp(u * math.log(hi / lo))).astype(np.float32)


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
