"""Pins of the multigrid metric coarsening (oracle/mg_oracle.py), CPU only."""
import numpy as np
import pytest

import mg_oracle as mo  # oracle/mg_oracle.py


def _fields(rng, nx, ny, nz):
    m = {k: rng.standard_normal((ny, nz, nx)) for k in mo.CELL}
    m.update({k: rng.standard_normal((ny, nz, nx + 1)) for k in mo.U_FACE})
    m.update({k: rng.standard_normal((ny + 1, nz, nx)) for k in mo.V_FACE})
    m["terrain"] = rng.standard_normal((ny, nx))
    return m, rng.standard_normal((ny, nz, nx))


def _cell_set(c, n):
    """Fine cells under coarse cell c (natural indices)."""
    return [i for i in (2 * c, 2 * c + 1) if i < n]


@pytest.mark.parametrize("nx,ny", [(1, 1), (2, 3), (5, 4), (7, 7), (6, 1)])
def test_coarsening_matches_the_index_sets(nx, ny):
    rng = np.random.default_rng(nx * 10 + ny)
    nz = 3
    m, alpha = _fields(rng, nx, ny, nz)
    got, a_c = mo.coarsen_metrics(m, alpha)
    nxc, nyc = mo.coarse_size(nx), mo.coarse_size(ny)
    for k in mo.CELL + ("alpha",):
        src = alpha if k == "alpha" else m[k]
        out = a_c if k == "alpha" else got[k]
        assert out.shape == (nyc, nz, nxc)
        for J in range(nyc):
            for I in range(nxc):
                cells = [(j, i) for j in _cell_set(J, ny) for i in _cell_set(I, nx)]
                want = np.mean([src[j, :, i] for j, i in cells], axis=0)
                assert np.allclose(out[J, :, I], want, rtol=1e-14, atol=1e-14)
    for k in mo.U_FACE:
        assert got[k].shape == (nyc, nz, nxc + 1)
        for J in range(nyc):
            for a in range(nxc + 1):
                want = np.mean([m[k][j, :, min(2 * a, nx)] for j in _cell_set(J, ny)], axis=0)
                assert np.allclose(got[k][J, :, a], want, rtol=1e-14, atol=1e-14)
    for k in mo.V_FACE:
        assert got[k].shape == (nyc + 1, nz, nxc)
        for b in range(nyc + 1):
            for I in range(nxc):
                want = np.mean([m[k][min(2 * b, ny), :, i] for i in _cell_set(I, nx)], axis=0)
                assert np.allclose(got[k][b, :, I], want, rtol=1e-14, atol=1e-14)
    for J in range(nyc):
        for I in range(nxc):
            want = np.mean([m["terrain"][j, i] for j in _cell_set(J, ny) for i in _cell_set(I, nx)])
            assert abs(got["terrain"][J, I] - want) <= 1e-14


def test_constants_stay_and_linear_fields_keep_their_mean_position():
    nx, ny, nz = 9, 6, 4
    m = {k: np.full((ny, nz, nx), 2.5) for k in mo.CELL}
    m.update({k: np.full((ny, nz, nx + 1), -1.0) for k in mo.U_FACE})
    m.update({k: np.full((ny + 1, nz, nx), 3.0) for k in mo.V_FACE})
    xi = np.arange(nx, dtype=float)
    m["terrain"] = np.tile(7.0 * xi, (ny, 1))
    got, a_c = mo.coarsen_metrics(m, np.ones((ny, nz, nx)))
    for k in mo.CELL:
        assert np.all(got[k] == 2.5)
    assert np.all(got["jacobian_u"] == -1.0) and np.all(got["dzdy_v"] == 3.0) and np.all(a_c == 1.0)
    # cell-centred means of a linear field are its value at the mean position
    pos = np.array([np.mean(_cell_set(I, nx)) for I in range(mo.coarse_size(nx))])
    assert np.allclose(got["terrain"], np.tile(7.0 * pos, (mo.coarse_size(ny), 1)), rtol=0, atol=1e-12)
