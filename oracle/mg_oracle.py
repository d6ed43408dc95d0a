"""Aggregation transfers of the solver's semi-coarsening multigrid, in numpy.

Test infrastructure only: imported by ``tests/`` -- never by the product
package.

The MIXED_MG inner preconditioner (no counterpart in the reference, whose
preconditioner is Jacobi-Richardson, wind_iterative.F90:1268-1315) builds
each coarse operator as a Galerkin product R A P: P copies the value of coarse
column (I, J) to the fine columns 2I-1, 2I and 2J-1, 2J (one of each at an odd
edge), every level keeps all nz+2 rows of a column, and R = P^T. This module
restates P, R and the application of a 15-point coefficient set on the
oracle's row layout (ny+2, nz+2, nx+2) (wind_oracle.c), so that R A P e can be
formed with the oracle's FP64 spmv (orc_spmv, wind_iterative.F90:813-1027) and
set against the coarse coefficients the solver reports.

Pins (tests/test_mg_oracle.py): P is checked against a brute-force loop over
the aggregate definition on tiny grids, R against <R y, e> = <y, P e>, and
apply15 against the oracle's own probed spmv (wind_iterative.F90:888-904) on
the 15 arrays of a probed operator.
"""
from __future__ import annotations

import numpy as np

# (di, dk, dj) of the stencil entries A..O (spmv, wind_iterative.F90:888-904)
OFFSETS = ((0, 0, 0), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0), (0, 0, 1), (0, 0, -1),
           (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (0, 1, 1), (0, 1, -1), (0, -1, 1), (0, -1, -1))


def coarse_size(n: int) -> int:
    return (n + 1) // 2


def cmap(nf: int) -> np.ndarray:
    """Fine index 0..nf+1 -> coarse index (lateral ring to lateral ring)."""
    idx = np.empty(nf + 2, dtype=np.int64)
    idx[0] = 0
    idx[nf + 1] = coarse_size(nf) + 1
    idx[1:nf + 1] = (np.arange(1, nf + 1) + 1) // 2
    return idx


def prolong(ec: np.ndarray, nxf: int, nyf: int) -> np.ndarray:
    """P e_c: coarse rows (nyc+2, mz, nxc+2) -> fine rows (nyf+2, mz, nxf+2),
    every fine interior row taking its aggregate's value, lateral rows 0."""
    ci, cj = cmap(nxf), cmap(nyf)
    out = ec[cj][:, :, ci].copy()
    out[0] = 0.0
    out[-1] = 0.0
    out[:, :, 0] = 0.0
    out[:, :, -1] = 0.0
    return out


def restrict(rf: np.ndarray, nxf: int, nyf: int) -> np.ndarray:
    """R r_f = P^T r_f: every coarse interior row sums its aggregate's rows;
    lateral coarse rows 0."""
    nxc, nyc = coarse_size(nxf), coarse_size(nyf)
    mz = rf.shape[1]
    out = np.zeros((nyc + 2, mz, nxc + 2))
    ci, cj = cmap(nxf), cmap(nyf)
    J = cj[1:nyf + 1][:, None, None]
    K = np.arange(mz)[None, :, None]
    I = ci[1:nxf + 1][None, None, :]
    np.add.at(out, (J, K, I), rf[1:nyf + 1, :, 1:nxf + 1])
    return out


def apply15(coef: np.ndarray, e: np.ndarray) -> np.ndarray:
    """(A e) on interior columns for 15 coefficient arrays in the row layout,
    coef[0] the row sum and coef[1..14] the couplings B..O: A e = s e +
    sum_q a_q (e_q - e). Lateral rows are identity. Couplings that would reach
    beyond the ghost planes multiply zero."""
    ny2, mz, nx2 = e.shape
    pad = np.zeros((ny2, mz + 2, nx2))
    pad[:, 1:mz + 1, :] = e
    y = e.copy()
    inner = (slice(1, ny2 - 1), slice(0, mz), slice(1, nx2 - 1))
    ec = e[inner]
    acc = coef[0][inner] * ec
    for q in range(1, 15):
        di, dk, dj = OFFSETS[q]
        nb = pad[1 + dj:ny2 - 1 + dj, 1 + dk:mz + 1 + dk, 1 + di:nx2 - 1 + di]
        acc = acc + coef[q][inner] * (nb - ec)
    y[inner] = acc
    return y


def solver_level_to_rows(coef_kji: np.ndarray) -> np.ndarray:
    """The solver's level coefficients [15][k][j][i] -> [15][j][k][i] (row layout)."""
    return np.ascontiguousarray(coef_kji.transpose(0, 2, 1, 3), dtype=np.float64)
