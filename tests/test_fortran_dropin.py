"""The Fortran host programs drive the solver through the ISO_C_BINDING
interface as HICAR would: tests/fortran/test_drop_in.F90 through the
array-level module, tests/fortran/test_domain_adapter.F90 through the
`wind_iterative` adapter with the reference's domain_t / options_t
signatures (against stub HICAR modules), two nests and the nest cache, and
tests/fortran/test_probe_mapfac.F90 through the adapter's probing API
(calibrate_projection_operator's 27-colour loop, wind.F90:1010-1099) and map
factors, whose map-factor solve is checked here against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "test_drop_in")
ADAPTER = os.path.join(ROOT, "tests", "fortran", "_build", "test_domain_adapter")
PROBE = os.path.join(ROOT, "tests", "fortran", "_build", "test_probe_mapfac")


def test_fortran_program_is_built():
    assert os.path.exists(EXE), "run __graft_entry__.build() (tests/fortran/Makefile)"
    assert os.path.exists(ADAPTER), "run __graft_entry__.build() (tests/fortran/Makefile)"
    assert os.path.exists(PROBE), "run __graft_entry__.build() (tests/fortran/Makefile)"


@pytest.mark.gpu
def test_fortran_drop_in_reduces_divergence():
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "PASS" in out.stdout


@pytest.mark.gpu
def test_fortran_domain_adapter_two_nests():
    out = subprocess.run([ADAPTER], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "PASS" in out.stdout and "FAIL" not in out.stdout
    # nest 2 runs verbose: the reference's per-solve timing breakdown, once
    assert out.stdout.count("per-solve timing breakdown") == 1
    spmv = [ln for ln in out.stdout.splitlines() if ln.strip().startswith("SpMV")]
    assert len(spmv) == 1 and float(spmv[0].split(":")[1].split()[0]) > 0


def _read_dump(path):
    """The stream file of test_probe_mapfac.F90: Fortran (i,k,j) arrays, read
    back as C-order (j,k,i) numpy arrays."""
    raw = open(path, "rb").read()
    nx, ny, nz = np.frombuffer(raw, np.int32, 3, 0)
    dx = float(np.frombuffer(raw, np.float32, 1, 12)[0])
    off = 16
    nx, ny, nz = int(nx), int(ny), int(nz)

    def take(shape, dtype=np.float32):
        nonlocal off
        n = int(np.prod(shape))
        a = np.frombuffer(raw, dtype, n, off).reshape(shape)
        off += n * np.dtype(dtype).itemsize
        return a.astype(np.float64)

    mass, ug, vg = (ny, nz, nx), (ny, nz, nx + 1), (ny + 1, nz, nx)
    m = {}
    for k, shp in (("dzdx", mass), ("dzdy", mass), ("jacobian", mass), ("jacobian_u", ug), ("jacobian_v", vg),
                   ("jacobian_w", mass), ("dzdx_u", ug), ("dzdy_v", vg), ("advection_dz", mass),
                   ("terrain", (ny, nx))):
        m[k] = take(shp)
    mf = {k: take(shp) for k, shp in (("mx_u", (ny, nx + 1)), ("my_u", (ny, nx + 1)), ("my_v", (ny + 1, nx)),
                                      ("mx_v", (ny + 1, nx)), ("mxy", (ny, nx)))}
    f = {k: take(shp) for k, shp in (("alpha", mass), ("rho", mass), ("u0", ug), ("v0", vg), ("w0", mass),
                                     ("div", mass))}
    status, iters = (int(q) for q in np.frombuffer(raw, np.int32, 2, off))
    off += 8
    f.update({k: take(shp) for k, shp in (("u", ug), ("v", vg), ("w", mass))})
    assert off == len(raw)
    return nx, ny, nz, dx, m, mf, f, status, iters


@pytest.mark.gpu
def test_fortran_probing_api_and_map_factors(tmp_path):
    """The five-call probing loop and probe_operator agree (in the program),
    and the adapter's map-factor divergence, solve and correction on FP32
    fields equal the oracle's FP64 results rounded to FP32."""
    from oracle import Oracle
    dump = tmp_path / "mapfac.bin"
    out = subprocess.run([PROBE, str(dump)], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "PASS" in out.stdout and "FAIL" not in out.stdout
    nx, ny, nz, dx, m, mf, f, status, iters = _read_dump(dump)
    o = Oracle(nx, ny, nz, dx)
    o.set_metrics(m)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)  # noqa: E731
    div = o.divergence(f["u0"], f["v0"], f["w0"], rho=f["rho"], adv_den=True, use_dqdt=True,
                       mf_my_u=mf["my_u"], mf_mx_v=mf["mx_v"], mf_mxy=mf["mxy"])
    assert np.array_equal(f32(div), f["div"])
    r = o.solve(f["alpha"], f["div"], tol_rel=1e-9, max_iters=5000)
    assert status == 0 and r["status"] == 0 and r["iters"] == iters
    u, v, w = o.correct(f["u0"], f["v0"], f["w0"], rho=f["rho"], mf_mx_u=mf["mx_u"], mf_my_v=mf["my_v"],
                        adv_den=True)
    for got, want in ((f["u"], u), (f["v"], v), (f["w"], w)):
        assert np.array_equal(got, f32(want))
    # and the map factors matter: without them the corrected u is elsewhere
    u1, _, _ = o.correct(f["u0"], f["v0"], f["w0"], rho=f["rho"], adv_den=True)
    assert np.abs(f32(u1) - f["u"]).max() > 1e3 * np.abs(f32(u) - f["u"]).max() + 1e-4
