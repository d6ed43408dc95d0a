"""Metric coarsening of the solver's semi-coarsening multigrid, in numpy.

Test infrastructure only: imported by ``tests/`` -- never by the product
package.

The MIXED_MG inner preconditioner (no counterpart in the reference, whose
preconditioner is Jacobi-Richardson, wind_iterative.F90:1268-1315) builds each
coarse level by rediscretising the analytic operator at twice the spacing
from averaged metrics. Coarse column c covers fine columns 2c and 2c + 1
(natural, 0-based indices; one column at an odd edge); the levels of a column
are kept. This module restates the averaging on the reference's natural
layouts (oracle/oracle.py: mass (ny, nz, nx), u faces (ny, nz, nx + 1), v
faces (ny + 1, nz, nx), 2-D (ny, nx)), so that an oracle built at the coarse
size and dx from these metrics assembles the operator the solver should hold
on that level.

Pins (tests/test_mg_oracle.py): every field against a brute-force loop over
the index sets the averaging is defined by, odd sizes included; constants stay
constant; a field linear in the cell index keeps its value at the mean
position.
"""
from __future__ import annotations

import numpy as np

CELL = ("dzdx", "dzdy", "jacobian", "jacobian_w", "advection_dz")
U_FACE = ("jacobian_u", "dzdx_u")
V_FACE = ("jacobian_v", "dzdy_v")


def coarse_size(n: int) -> int:
    return (n + 1) // 2


def _cells(a: np.ndarray, axis: int) -> np.ndarray:
    """Mean over the fine cells 2c, 2c+1 of every coarse cell c along `axis`."""
    n = a.shape[axis]
    nc = coarse_size(n)
    lo = np.take(a, np.arange(0, 2 * nc, 2), axis=axis)
    hi_idx = np.minimum(np.arange(1, 2 * nc, 2), n - 1)
    hi = np.take(a, hi_idx, axis=axis)
    two = (np.arange(1, 2 * nc, 2) <= n - 1).astype(float)
    shape = [1] * a.ndim
    shape[axis] = nc
    two = two.reshape(shape)
    return (lo + hi * two) / (1.0 + two)


def _faces(a: np.ndarray, axis: int) -> np.ndarray:
    """The fine face under every coarse face a_c: fine face min(2 a_c, n)
    (n + 1 faces along `axis` on the fine level, nc + 1 on the coarse)."""
    n = a.shape[axis] - 1
    nc = coarse_size(n)
    return np.take(a, np.minimum(2 * np.arange(nc + 1), n), axis=axis)


def coarsen_metrics(m: dict, alpha: np.ndarray):
    """Averaged metrics and alpha of the next coarser level (natural layouts)."""
    out = {}
    for k in CELL:
        out[k] = _cells(_cells(np.asarray(m[k], dtype=np.float64), 2), 0)
    for k in U_FACE:
        out[k] = _cells(_faces(np.asarray(m[k], dtype=np.float64), 2), 0)
    for k in V_FACE:
        out[k] = _faces(_cells(np.asarray(m[k], dtype=np.float64), 2), 0)
    out["terrain"] = _cells(_cells(np.asarray(m["terrain"], dtype=np.float64), 1), 0)
    return out, _cells(_cells(np.asarray(alpha, dtype=np.float64), 2), 0)
