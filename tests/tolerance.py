"""The tolerance between exact FP32 filter weights and the reference's 8-bit fixed-point texture
weights (source/granite/TrajectoryIntegrator.icc:22-40: tex3D / tex2D linear filtering), as a
bound computed from the field alone.

Per sample: every linear interpolation weight of the hardware filter differs from the exact one
by at most delta (2^-9 when the weight is rounded to 1/256, 2^-8 when truncated). A multilinear
interpolation is a chain of convex combinations, one axis after the other, so a weight error on
axis a changes the sample by at most delta * D_a, with D_a the largest difference between two
nodes adjacent along a (any component, either time slice); later axes only average that change.
Hence |u_q8(p) - u(p)| <= eps = delta * sum_a D_a at every point, and the time lerp (the same in
both) keeps the bound.

Per trajectory: the exactly interpolated field is Lipschitz in the max norm with
L = sum_a D_a / h_a. Grönwall's argument on one step of size dt (z = dt L):
  Euler  |e_{k+1}| <= (1 + z) |e_k| + dt eps
  RK4    |e_{k+1}| <= A |e_k| + dt eps g,  A = 1 + z + z^2/2 + z^3/6 + z^4/24,
         g = 1 + z/2 + z^2/6 + z^3/24  (stage k_i perturbed by eps plus L times the error of the
         stage position it is evaluated at, weighted 1, 2, 2, 1 over 6)
so |e_n| <= (dt eps g + r) (A^n - 1) / (A - 1), where r bounds what rounding adds per step: two
FP32 runs of the same step (or a run whose compiler contracts a * b + c into one FMA, as the
reference's nvcc build does in its stage sums) differ by a few ulp of the position; r counts 16
ulp of the largest coordinate, and it propagates through the same recursion."""
from __future__ import annotations

import numpy as np

DELTA = {"Quantized8": 2.0 ** -9, "Quantized8Truncated": 2.0 ** -8}


def node_differences(wind: np.ndarray, slices) -> np.ndarray:
    """D_a per spatial axis a (x first) over the given time slices of wind (S, C, [nz,] ny, nx):
    the largest |difference| between nodes adjacent along a, over every component."""
    w = np.asarray(wind, np.float64)[list(slices)]
    dim = w.ndim - 2
    out = []
    for a in range(dim):
        ax = w.ndim - 1 - a  # x is the last array axis
        out.append(float(np.max(np.abs(np.diff(w, axis=ax)))))
    return np.array(out)


def position_bound(D, h, dt, steps, delta, rk4, scale):
    """Bound on max_i |x_i^q8(k) - x_i(k)|_inf for k = 0..steps (array of steps + 1 values);
    scale: the largest |coordinate| (for the rounding slack)."""
    D, h = np.asarray(D, np.float64), np.asarray(h, np.float64)
    eps = delta * D.sum()
    L = float((D / h).sum())
    z = dt * L
    if rk4:
        A = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
        g = 1 + z / 2 + z * z / 6 + z ** 3 / 24
    else:
        A, g = 1 + z, 1.0
    k = np.arange(steps + 1, dtype=np.float64)
    grow = k if A == 1 else (A ** k - 1) / (A - 1)
    r = 16 * float(np.spacing(np.float32(scale)))
    return (dt * eps * g + r) * grow
