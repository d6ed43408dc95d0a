"""ctypes front end of the CPU oracle (``oracle/wind_oracle.c``).

Test infrastructure only: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` -- never by the product package.

Array conventions (numpy, C order):
  mass fields  (ny, nz, nx)      == Fortran (i,k,j) over (nx, nz, ny)
  u fields     (ny, nz, nx+1)
  v fields     (ny+1, nz, nx)
  2-D fields   (ny, nx)
  rows         (ny+2, nz+2, nx+2) == solver rows (i,k,j) over 0..mx-1 etc.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None
_variants = {}

COEF_NAMES = "ABCDEFGHIJKLMNO"
_dp = C.POINTER(C.c_double)


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load(path=None):
    """The parity build (oracle/liboracle.so), or, with `path`, another build
    of the same source (the CPU-baseline builds of the Makefile's `baseline`)."""
    global _lib
    if path is not None:
        if path not in _variants:
            _variants[path] = _bind(C.CDLL(path))
        return _variants[path]
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    _lib = _bind(C.CDLL(_LIB_PATH))
    return _lib


def _bind(lib):
    v = C.c_void_p
    i = C.c_int
    d = C.c_double
    sig = {
        "orc_create": (v, [i, i, i, d]),
        "orc_free": (None, [v]),
        "orc_rows": (C.c_long, [v]),
        "orc_set_metrics": (None, [v] + [_dp] * 10),
        "orc_set_alpha": (None, [v, _dp]),
        "orc_update_coefs": (None, [v]),
        "orc_assemble": (None, [v]),
        "orc_spmv": (None, [v, _dp, _dp]),
        "orc_extract_diagonal": (None, [v]),
        "orc_compute_rhs": (None, [v, _dp]),
        "orc_apply_precond": (None, [v, i, _dp, _dp]),
        "orc_bicgstab": (i, [v, i, d, d, C.POINTER(i), C.POINTER(i), _dp, _dp, _dp, i]),
        "orc_solve": (i, [v, _dp, _dp, i, d, d, C.POINTER(i), C.POINTER(i), _dp, _dp, C.POINTER(i)]),
        "orc_set_nsweeps": (None, [v, i]),
        "orc_get_lambda": (None, [v, _dp]),
        "orc_set_lambda": (None, [v, _dp]),
        "orc_get_rows": (None, [v, i, _dp]),
        "orc_get_coef": (None, [v, i, _dp]),
        "orc_set_coef": (None, [v, i, _dp]),
        "orc_set_probed": (None, [v, i]),
        "orc_get_planes": (None, [v, _dp, _dp, _dp]),
        "orc_correct": (None, [v, _dp, _dp, _dp, _dp, _dp, _dp, i]),
        "orc_divergence": (None, [v, _dp, _dp, _dp, _dp, _dp, _dp, _dp, i, i, i, _dp]),
        "orc_balance_uvw": (None, [v, _dp, _dp, _dp, _dp, i, i]),
        "orc_w_real": (None, [v, _dp, _dp, _dp, _dp]),
        "orc_idealized_wgrid": (None, [v, _dp, _dp, _dp, _dp]),
        "orc_probe": (d, [v, _dp, _dp, _dp, _dp, _dp, _dp, i]),
        "orc_canon_reduce": (d, [_dp, i, i]),
        "orc_dot": (d, [v, _dp, _dp]),
        "orc_num_threads": (i, []),
        "orc_plant": (None, [i]),
        "orc_set_faithful": (None, [v, i]),
        "orc_set_threads": (None, [i]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def canon_reduce(colsum: np.ndarray) -> float:
    """Canonical sum of per-column values given as a (my, mx) array."""
    cs = _f64(colsum)
    my, mx = cs.shape
    return _load().orc_canon_reduce(_p(cs), mx, my)


def num_threads() -> int:
    return _load().orc_num_threads()


def plant(code: int) -> None:
    """Make the oracle commit one planted mistake (codes in wind_oracle.c,
    0 = none). Tests only: they show that a pin fails on it."""
    _load().orc_plant(int(code))


class Oracle:
    """The reference path restated on the CPU, FP64 throughout."""

    def __init__(self, nx: int, ny: int, nz: int, dx: float, faithful: bool = False, lib_path: str | None = None):
        """faithful: round to FP32 where the reference computes in `real`
        (module arrays, assembly, rhs product, correction); see
        orc_set_faithful in wind_oracle.c. lib_path: another build of the
        oracle (the CPU-baseline builds), default the parity build."""
        self.lib = _load(lib_path)
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, float(dx)
        self.h = self.lib.orc_create(nx, ny, nz, float(dx))
        if not self.h:
            raise ValueError("oracle: grid too small (need nx, ny >= 3 and nz >= 2)")
        self.faithful = bool(faithful)
        if faithful:
            self.lib.orc_set_faithful(self.h, 1)
        self.rows_shape = (ny + 2, nz + 2, nx + 2)

    def close(self):
        if self.h:
            self.lib.orc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- set-up
    def set_metrics(self, m: dict) -> None:
        keep = [_f64(m[k]) for k in ("dzdx", "dzdy", "jacobian", "jacobian_u", "jacobian_v",
                                     "jacobian_w", "dzdx_u", "dzdy_v", "advection_dz", "terrain")]
        self.lib.orc_set_metrics(self.h, *[_p(a) for a in keep])

    def set_alpha(self, alpha):
        a = _f64(alpha)
        self.lib.orc_set_alpha(self.h, _p(a))

    def assemble(self):
        self.lib.orc_assemble(self.h)

    def update_coefs(self):
        self.lib.orc_update_coefs(self.h)

    def extract_diagonal(self):
        self.lib.orc_extract_diagonal(self.h)

    def compute_rhs(self, div):
        d = _f64(div)
        self.lib.orc_compute_rhs(self.h, _p(d))

    def set_nsweeps(self, n: int):
        self.lib.orc_set_nsweeps(self.h, n)

    def set_probed(self, probed: bool):
        self.lib.orc_set_probed(self.h, 1 if probed else 0)

    # ---- operators
    def spmv(self, x):
        x = _f64(x).reshape(self.rows_shape)
        y = np.zeros(self.rows_shape)
        self.lib.orc_spmv(self.h, _p(x), _p(y))
        return y

    def apply_precond(self, nsweeps: int, vin):
        vin = _f64(vin).reshape(self.rows_shape)
        out = np.zeros(self.rows_shape)
        self.lib.orc_apply_precond(self.h, nsweeps, _p(vin), _p(out))
        return out

    def dot(self, a, b) -> float:
        a = _f64(a).reshape(self.rows_shape)
        b = _f64(b).reshape(self.rows_shape)
        return self.lib.orc_dot(self.h, _p(a), _p(b))

    def bicgstab(self, max_iters=1500, tol_rel=1e-5, tol_abs=1e-10, trace_cap=0):
        st, it = C.c_int(), C.c_int()
        r0, rf = C.c_double(), C.c_double()
        tr = np.zeros(max(1, 5 * trace_cap))
        self.lib.orc_bicgstab(self.h, max_iters, tol_rel, tol_abs, C.byref(st), C.byref(it),
                              C.byref(r0), C.byref(rf), _p(tr), trace_cap)
        out = dict(status=st.value, iters=it.value, res0=r0.value, res_final=rf.value)
        if trace_cap:
            out["trace"] = tr.reshape(-1, 5)[: min(trace_cap, it.value)]
        return out

    def solve(self, alpha, div, max_iters=1500, tol_rel=1e-5, tol_abs=1e-10):
        """calc_iter_winds without the correction: assembly/refresh, diagonal,
        rhs, BiCGStab from the current lambda, retry ladder."""
        a, d = _f64(alpha), _f64(div)
        st, it, ns = C.c_int(), C.c_int(), C.c_int()
        r0, rf = C.c_double(), C.c_double()
        rc = self.lib.orc_solve(self.h, _p(a), _p(d), max_iters, tol_rel, tol_abs, C.byref(st),
                                C.byref(it), C.byref(r0), C.byref(rf), C.byref(ns))
        return dict(rc=rc, status=st.value, iters=it.value, res0=r0.value, res_final=rf.value,
                    nsweeps=ns.value)

    # ---- state access
    def get_lambda(self):
        out = np.zeros(self.rows_shape)
        self.lib.orc_get_lambda(self.h, _p(out))
        return out

    def set_lambda(self, lam):
        lam = _f64(lam).reshape(self.rows_shape)
        self.lib.orc_set_lambda(self.h, _p(lam))

    def rows(self, which: str):
        out = np.zeros(self.rows_shape)
        self.lib.orc_get_rows(self.h, {"rhs": 0, "dinv": 1, "r": 2, "x": 3}[which], _p(out))
        return out

    def coef(self, name: str):
        out = np.zeros((self.ny, self.nz, self.nx))
        self.lib.orc_get_coef(self.h, COEF_NAMES.index(name), _p(out))
        return out

    def set_coef(self, name: str, arr):
        a = _f64(arr)
        self.lib.orc_set_coef(self.h, COEF_NAMES.index(name), _p(a))

    def planes(self):
        dz_if = np.zeros((self.ny, self.nz + 1, self.nx))
        sx = np.zeros((self.ny, self.nx))
        sy = np.zeros((self.ny, self.nx))
        self.lib.orc_get_planes(self.h, _p(dz_if), _p(sx), _p(sy))
        return dz_if, sx, sy

    # ---- winds
    def correct(self, u, v, w, rho=None, mf_mx_u=None, mf_my_v=None, adv_den=False):
        """In-place u/v/w correction from the current lambda; returns copies."""
        u, v, w = _f64(u).copy(), _f64(v).copy(), _f64(w).copy()
        rho, mx, my = _f64(rho), _f64(mf_mx_u), _f64(mf_my_v)
        self.lib.orc_correct(self.h, _p(u), _p(v), _p(w), _p(rho), _p(mx), _p(my), int(adv_den))
        return u, v, w

    def divergence(self, u, v, w, rho=None, adv_den=True, horz_only=False, use_dqdt=True,
                   mf_my_u=None, mf_mx_v=None, mf_mxy=None):
        u, v, w = _f64(u), _f64(v), _f64(w)
        rho = _f64(rho)
        a, b, c = _f64(mf_my_u), _f64(mf_mx_v), _f64(mf_mxy)
        div = np.zeros((self.ny, self.nz, self.nx))
        self.lib.orc_divergence(self.h, _p(u), _p(v), _p(w), _p(rho), _p(a), _p(b), _p(c),
                                int(adv_den), int(horz_only), int(use_dqdt), _p(div))
        return div

    def balance_uvw(self, u, v, w, rho=None, adv_den=True, use_dqdt=True):
        """balance_uvw + calc_w (wind.F90:123-234): a new w (mass layout)."""
        u, v, rho = _f64(u), _f64(v), _f64(rho)
        out = np.array(_f64(w), dtype=np.float64, copy=True)
        self.lib.orc_balance_uvw(self.h, _p(u), _p(v), _p(out), _p(rho), int(adv_den), int(use_dqdt))
        return out

    def w_real(self, u, v, w_grid=None):
        """calc_w_real (wind.F90:1175-1227)."""
        u, v, wg = _f64(u), _f64(v), _f64(w_grid)
        out = np.zeros((self.ny, self.nz, self.nx))
        self.lib.orc_w_real(self.h, _p(u), _p(v), _p(wg), _p(out))
        return out

    def idealized_wgrid(self, u, v, w_real):
        """calc_idealized_wgrid (wind.F90:940-998)."""
        u, v, wr = _f64(u), _f64(v), _f64(w_real)
        out = np.zeros((self.ny, self.nz, self.nx))
        self.lib.orc_idealized_wgrid(self.h, _p(u), _p(v), _p(wr), _p(out))
        return out

    def probe(self, rho=None, adv_den=False, mf_mx_u=None, mf_my_v=None, mf_my_u=None, mf_mx_v=None, mf_mxy=None):
        """calibrate_projection_operator (wind.F90:1010-1099); map factors as in
        correct (G) and divergence (D), None = 1."""
        r = _f64(rho)
        mfs = [_f64(a) for a in (mf_mx_u, mf_my_v, mf_my_u, mf_mx_v, mf_mxy)]
        return self.lib.orc_probe(self.h, _p(r), *[_p(a) for a in mfs], int(adv_den))
