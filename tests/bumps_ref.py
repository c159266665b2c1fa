"""Host restatement of the seeded initial condition of the precision benchmark
(diagonal discontinuity plus 16 Gaussian bumps in rho and p), test
infrastructure: it checks the device's initial state (e2d_init_problem with
E2D_PROBLEM_BUMPS), it is not used by the library.

The problem (diag): x + y < 1 with x = (i-2)/nx, y = (j-2)/ny (0-based cell
index, ghosts included, the reference's tmp of serial_ref/m_run.f90:539 with
the diagonal moved to 1): rho = 1, p = 1, else rho = 0.125, p = 0.1; u = v = 0,
E = p / (gamma - 1).  The bumps: 16 draws (centre x, centre y, amplitude on
rho in [0.1, 2], amplitude on p in [0.1, 2], sigma in [0.02, 0.1]) of
SplitMix64 (Steele, Lea and Flood, OOPSLA 2014) uniforms u = (z >> 11) 2^-53,
and rho += sum_k A_rho,k w_k, p += sum_k A_p,k w_k with
w_k = exp(-|x_c - c_k|^2 / (2 sigma_k^2)) at the cell centre
x_c = ((i-2) + 1/2) / nx, summed over k in order in float64.
"""
from __future__ import annotations

import numpy as np

_M64 = (1 << 64) - 1


def splitmix64(state: int):
    """(next state, output) of SplitMix64."""
    state = (state + 0x9E3779B97F4A7C15) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return state, z ^ (z >> 31)


def bump_draws(seed: int) -> np.ndarray:
    """(16, 5): centre x, centre y, amplitude on rho, amplitude on p, sigma."""
    s = seed & _M64
    out = np.zeros((16, 5))
    lo = np.array([0.0, 0.0, 0.1, 0.1, 0.02])
    width = np.array([1.0, 1.0, 1.9, 1.9, 0.08])
    for k in range(16):
        for m in range(5):
            s, z = splitmix64(s)
            out[k, m] = lo[m] + width[m] * ((z >> 11) * 2.0 ** -53)
    return out


def bumps_state(nx: int, ny: int, gamma0: float, seed: int, arith=np.float64, store=np.float64) -> np.ndarray:
    """Initial state u(isize, jsize, 4): the diagonal test, the additions and E
    in the arithmetic type `arith` (float64, or float32 for the FP32 run), the
    bump sums in float64, the result rounded to the storage type `store`;
    returned as float64."""
    b = bump_draws(seed)
    i = np.arange(nx + 4)[:, None]
    j = np.arange(ny + 4)[None, :]
    T = np.dtype(arith).type
    x = (i - 2).astype(arith) / T(nx)
    y = (j - 2).astype(arith) / T(ny)
    low = (x + y) < T(1)
    rho = np.where(low, T(1), T(0.125)).astype(arith)
    p = np.where(low, T(1), T(0.1)).astype(arith)
    xc = ((i - 2) + 0.5) / nx
    yc = ((j - 2) + 0.5) / ny
    dr = np.zeros((nx + 4, ny + 4))
    dp = np.zeros((nx + 4, ny + 4))
    for k in range(16):
        ddx, ddy, sg = xc - b[k, 0], yc - b[k, 1], b[k, 4]
        w = np.exp(-(ddx * ddx + ddy * ddy) / (2.0 * sg * sg))
        dr = dr + b[k, 2] * w
        dp = dp + b[k, 3] * w
    rho = (rho + dr.astype(arith)).astype(arith)
    p = (p + dp.astype(arith)).astype(arith)
    e = (p / (T(gamma0) - T(1))).astype(arith)
    u = np.zeros((nx + 4, ny + 4, 4))
    u[..., 0] = rho.astype(store)
    u[..., 1] = e.astype(store)
    return u
