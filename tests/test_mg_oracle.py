"""Pins of the multigrid transfer restatement (oracle/mg_oracle.py), CPU only."""
import numpy as np
import pytest

import mg_oracle as mo  # oracle/mg_oracle.py
from conftest import oracle_for, small_case


@pytest.mark.parametrize("nx,ny", [(1, 1), (2, 3), (5, 4), (7, 7), (6, 1)])
def test_prolong_matches_aggregate_definition(nx, ny):
    rng = np.random.default_rng(nx * 10 + ny)
    nxc, nyc, mz = mo.coarse_size(nx), mo.coarse_size(ny), 3
    ec = rng.standard_normal((nyc + 2, mz, nxc + 2))
    want = np.zeros((ny + 2, mz, nx + 2))
    for j in range(1, ny + 1):
        for i in range(1, nx + 1):
            # fine columns 2I-1, 2I belong to coarse column I
            I = next(I for I in range(1, nxc + 1) if i in (2 * I - 1, 2 * I))
            J = next(J for J in range(1, nyc + 1) if j in (2 * J - 1, 2 * J))
            want[j, :, i] = ec[J, :, I]
    assert np.array_equal(mo.prolong(ec, nx, ny), want)
    # every coarse interior column receives at least one fine column, at most four
    hits = np.zeros((nyc + 2, nxc + 2), dtype=int)
    ci, cj = mo.cmap(nx), mo.cmap(ny)
    for j in range(1, ny + 1):
        for i in range(1, nx + 1):
            hits[cj[j], ci[i]] += 1
    assert hits[1:-1, 1:-1].min() >= 1 and hits.max() <= 4 and hits[0].sum() == 0


@pytest.mark.parametrize("nx,ny", [(2, 3), (5, 4), (9, 6)])
def test_restrict_is_the_transpose_of_prolong(nx, ny):
    rng = np.random.default_rng(7)
    mz = 4
    nxc, nyc = mo.coarse_size(nx), mo.coarse_size(ny)
    y = rng.standard_normal((ny + 2, mz, nx + 2))
    e = rng.standard_normal((nyc + 2, mz, nxc + 2))
    e[0] = e[-1] = 0.0
    e[:, :, 0] = e[:, :, -1] = 0.0
    lhs = float(np.sum(mo.restrict(y, nx, ny) * e))
    rhs = float(np.sum(y * mo.prolong(e, nx, ny)))
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1.0)


def test_apply15_matches_the_oracle_probed_spmv():
    """apply15 on a random 15-array operator equals the oracle's probed spmv
    (wind_iterative.F90:888-904) on every interior row."""
    c = small_case("fractal")
    o = oracle_for(c)
    o.set_alpha(c.alpha)
    o.assemble()
    rng = np.random.default_rng(3)
    a = [rng.standard_normal((c.ny, c.nz, c.nx)) for _ in range(15)]
    o.set_probed(True)
    for q, name in enumerate("ABCDEFGHIJKLMNO"):
        o.set_coef(name, a[q])
    x = rng.standard_normal(o.rows_shape)
    want = o.spmv(x)
    coef = np.zeros((15,) + o.rows_shape)
    for q in range(1, 15):
        coef[q][1:-1, 1:-1, 1:-1] = a[q]
    coef[0][1:-1, 1:-1, 1:-1] = sum(a)  # row sum
    got = mo.apply15(coef, x)
    inner = (slice(1, -1), slice(1, -1), slice(1, -1))
    assert np.allclose(got[inner], want[inner], rtol=1e-12, atol=1e-12)
    o.close()
