"""Shared test helpers: build pygranite settings from a dict, run the product through its
pygranite-compatible API, and compare with the oracle."""
from __future__ import annotations

import numpy as np

from pygranite_amd import synthetic


def make_settings(pg, d: dict):
    s = pg.IntegratorSettings()
    enums = {
        "Space": pg.Space, "Integrator": pg.Integrator, "BorderMode": pg.BorderMode,
        "CurvatureMode": pg.CurvatureMode, "AbortMode": pg.AbortMode, "UpLiftMode": pg.UpLiftMode,
        "WindfieldMode": pg.WindfieldMode, "ConstantsMode": pg.ConstantsMode,
        "AdditionalVolumeMode": pg.AdditionalVolumeMode, "WindfieldStorage": pg.StorageType,
        "TextureWeights": pg.TextureWeights,
    }
    for k, v in d.items():
        if k in enums and isinstance(v, str):
            v = getattr(enums[k], v)
        setattr(s, k, v)
    return s


def run_product(pg, settings: dict, wind, seeds, *, uplift=None, constants=None, volumes=None,
                topography=None):
    """Runs one job through pygranite's TrajectoryIntegrator; returns a dict like oracle.run."""
    s = make_settings(pg, settings)
    if topography is not None:
        s.Topography = np.ascontiguousarray(topography, np.float32)
    Loader = synthetic.make_loader_class(pg.DataLoader)
    loader = Loader(wind, uplift=uplift, constants=constants, volumes=volumes)
    ts = pg.TrajectorySet(np.ascontiguousarray(seeds, np.float32))
    integ = pg.TrajectoryIntegrator(s, loader, ts)
    out = integ.compute()
    res = {
        "trajectories": out.trajectories(),
        "status": np.array([int(r.value) for r in out.abortReasons()], np.int32),
        "abort_step": out.abortSteps(),
        "total_curvature": out.totalCurvatures(),
        "curvatures": out.curvatures(),
        "lengths": out.individualLengths(),
        "compute": [out.computeInfo(i) for i in range(len(settings.get("AdditionalCompute", [])))],
        "set": out,
        "loader": loader,
    }
    if volumes:
        res["volumes"] = {k: out.volumeInfo(k) for k in volumes}
    return res


def assert_bitexact(gpu: dict, ref: dict, *, fields=("trajectories",), label=""):
    """Exact equality of float32 product outputs with the FP32 oracle's outputs."""
    for f in fields:
        a = np.asarray(gpu[f], np.float32)
        b = np.asarray(ref[f]).astype(np.float32)
        assert a.shape == b.shape, f"{label}{f}: shape {a.shape} vs {b.shape}"
        bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
        if bad.any():
            idx = np.argwhere(bad)[0]
            raise AssertionError(f"{label}{f}: {bad.sum()} of {a.size} differ; first at {idx}: "
                                 f"{a[tuple(idx)]!r} vs {b[tuple(idx)]!r}")
    s_gpu = np.asarray(gpu["status"])
    s_ref = np.asarray(ref["status"])
    assert np.array_equal(s_gpu, s_ref), f"{label}status differs at {np.flatnonzero(s_gpu != s_ref)[:10]}"
    st_gpu = np.asarray(gpu["abort_step"]).astype(np.int64)
    st_ref = np.asarray(ref["abort_step"]).astype(np.int64)
    assert np.array_equal(st_gpu, st_ref), f"{label}abort step differs at {np.flatnonzero(st_gpu != st_ref)[:10]}"


def domain_rel_error(a, b, extent):
    """max and RMS of |a - b| / extent over all saved positions (per axis extent)."""
    d = np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)) / np.asarray(extent, np.float64)
    d = d[np.isfinite(d)]
    return float(d.max()) if d.size else 0.0, float(np.sqrt(np.mean(d * d))) if d.size else 0.0
