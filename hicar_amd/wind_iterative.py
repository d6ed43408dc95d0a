"""Host-side mirror of HICAR's ``wind_iterative`` module over the HIP solver.

Same operations as the reference module (src/physics/wind_iterative.F90:39-41):
``init_iter_winds``, ``calc_iter_winds``, ``finalize_iter_winds`` and the operator
probing hooks, with the same argument meaning; the arithmetic runs in
libwindhip.so (hicar_amd/csrc). Arrays are numpy, C order == Fortran (i,k,j):
mass (ny, nz, nx), u (ny, nz, nx+1), v (ny+1, nz, nx), 2-D (ny, nx).

Differences from the reference, by design:
* errors raise ``WindHipError`` instead of printing and calling ``stop``
  (wind_iterative.F90:506-528);
* the plain path keeps lambda in FP64 through the correction (the reference
  rounds it to FP32, :1770-1778).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as nat

_METRIC_KEYS = ("dzdx", "dzdy", "jacobian", "jacobian_u", "jacobian_v", "jacobian_w",
                "dzdx_u", "dzdy_v", "advection_dz", "terrain")
_MAPFAC_KEYS = ("mapfac_mx_u", "mapfac_my_u", "mapfac_my_v", "mapfac_mx_v", "mapfac_mxy")


class _DevArr:
    """A device field argument: checks shape, dtype and contiguity of a tensor-like
    object and exposes its device address."""

    def __init__(self, t, shape, dtype, name):
        if not hasattr(t, "data_ptr"):
            raise TypeError(f"{name}: device fields need an object with data_ptr() (e.g. a GPU tensor)")
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name}: expected shape {shape}, got {tuple(t.shape)}")
        if getattr(t, "element_size", lambda: dtype.itemsize)() != dtype.itemsize:
            raise ValueError(f"{name}: expected {dtype} elements")
        if hasattr(t, "is_contiguous") and not t.is_contiguous():
            raise ValueError(f"{name}: must be contiguous")
        self.addr = int(t.data_ptr())


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, _DevArr):
        return C.c_void_p(a.addr)
    return C.c_void_p(a.ctypes.data)


class WindSolver:
    """One nest's solver state on the GPU (the reference's module state,
    wind_iterative.F90:42-205, held per nest as in hicar_cache_t :142-181)."""

    def __init__(self, nx: int, ny: int, nz: int, dx: float, dtype=np.float64, device_fields: bool = False):
        """device_fields: field arguments are device arrays (objects with
        ``data_ptr()``, e.g. torch tensors on the GPU) instead of numpy arrays."""
        self.lib = nat.load()
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, float(dx)
        self.dtype = np.dtype(dtype)
        self.device_fields = bool(device_fields)
        if self.dtype not in (np.dtype(np.float32), np.dtype(np.float64)):
            raise ValueError("dtype must be float32 (HICAR real) or float64")
        d = nat.Desc(nx=nx, ny=ny, nz=nz, dx=float(dx), ims=1, ime=nx, kms=1, kme=nz, jms=1, jme=ny,
                     field_dtype=self.dtype.itemsize,
                     field_mem=nat.WINDHIP_DEVICE if device_fields else nat.WINDHIP_HOST)
        h = C.c_void_p()
        nat.check(self.lib.windhip_create(C.byref(d), C.byref(h)))
        self.h = h
        self._keep = []

    # ----------------------------------------------------------------- helpers
    def _arr(self, a, shape, name):
        if a is None:
            return None
        if self.device_fields:
            return _DevArr(a, shape, self.dtype, name)
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.shape != shape:
            raise ValueError(f"{name}: expected shape {shape}, got {a.shape}")
        return a

    @property
    def mass_shape(self):
        return (self.ny, self.nz, self.nx)

    @property
    def u_shape(self):
        return (self.ny, self.nz, self.nx + 1)

    @property
    def v_shape(self):
        return (self.ny + 1, self.nz, self.nx)

    @property
    def rows_shape(self):
        return (self.ny + 2, self.nz + 2, self.nx + 2)

    # ------------------------------------------------------------------- API
    def set_metrics(self, m: dict) -> None:
        """init_module_vars (wind_iterative.F90:2274-2421)."""
        shapes = dict(dzdx=self.mass_shape, dzdy=self.mass_shape, jacobian=self.mass_shape,
                      jacobian_u=self.u_shape, jacobian_v=self.v_shape, jacobian_w=self.mass_shape,
                      dzdx_u=self.u_shape, dzdy_v=self.v_shape, advection_dz=self.mass_shape,
                      terrain=(self.ny, self.nx))
        arrs = [self._arr(m[k], shapes[k], k) for k in _METRIC_KEYS]
        mf_shapes = dict(mapfac_mx_u=(self.ny, self.nx + 1), mapfac_my_u=(self.ny, self.nx + 1),
                         mapfac_my_v=(self.ny + 1, self.nx), mapfac_mx_v=(self.ny + 1, self.nx),
                         mapfac_mxy=(self.ny, self.nx))
        mfs = [self._arr(m.get(k), mf_shapes[k], k) for k in _MAPFAC_KEYS]
        nat.check(self.lib.windhip_set_metrics(self.h, *[_ptr(a) for a in arrs + mfs]))

    def solve(self, alpha, div, **opts) -> dict:
        """The solve of calc_iter_winds (wind_iterative.F90:295-528)."""
        a = self._arr(alpha, self.mass_shape, "alpha")
        d = self._arr(div, self.mass_shape, "div")
        o = nat.default_opts(**opts)
        info = nat.SolveInfo()
        nat.check(self.lib.windhip_solve(self.h, _ptr(a), _ptr(d), C.byref(o), C.byref(info)))
        return info.as_dict()

    def apply_correction(self, u, v, w, density=None, adv_den=True):
        """calc_updated_winds (wind_iterative.F90:1742-2071); u, v, w updated in place."""
        if self.device_fields:
            uu, vv, ww = (self._arr(u, self.u_shape, "u"), self._arr(v, self.v_shape, "v"),
                          self._arr(w, self.mass_shape, "w"))
            rho = self._arr(density, self.mass_shape, "density")
            nat.check(self.lib.windhip_apply_correction(self.h, _ptr(uu), _ptr(vv), _ptr(ww), _ptr(rho),
                                                        int(bool(adv_den))))
            return
        for arr, shp, nm in ((u, self.u_shape, "u"), (v, self.v_shape, "v"), (w, self.mass_shape, "w")):
            if not (isinstance(arr, np.ndarray) and arr.dtype == self.dtype and arr.flags.c_contiguous
                    and arr.shape == shp):
                raise ValueError(f"{nm}: need a C-contiguous {self.dtype} array of shape {shp} (updated in place)")
        rho = self._arr(density, self.mass_shape, "density")
        nat.check(self.lib.windhip_apply_correction(self.h, _ptr(u), _ptr(v), _ptr(w), _ptr(rho),
                                                    int(bool(adv_den))))

    def calc_iter_winds(self, alpha, div, u, v, w, density=None, adv_den=True, **opts) -> dict:
        """calc_iter_winds(domain, alpha, div, adv_den) (wind_iterative.F90:295)."""
        info = self.solve(alpha, div, **opts)
        self.apply_correction(u, v, w, density, adv_den)
        return info

    def divergence(self, u, v, w, density=None, adv_den=True, horz_only=False, use_dqdt=True, out=None):
        """calc_divergence (wind.F90:250-523). With device fields `out` must be given."""
        uu = self._arr(u, self.u_shape, "u")
        vv = self._arr(v, self.v_shape, "v")
        ww = self._arr(w, self.mass_shape, "w")
        rho = self._arr(density, self.mass_shape, "density")
        if self.device_fields:
            if out is None:
                raise ValueError("divergence: pass `out` (a device array) when fields live on the device")
            nat.check(self.lib.windhip_divergence(self.h, _ptr(uu), _ptr(vv), _ptr(ww), _ptr(rho),
                                                  int(bool(adv_den)), int(bool(horz_only)), int(bool(use_dqdt)),
                                                  _ptr(self._arr(out, self.mass_shape, "out"))))
            return out
        out = np.zeros(self.mass_shape, dtype=self.dtype)
        nat.check(self.lib.windhip_divergence(self.h, _ptr(uu), _ptr(vv), _ptr(ww), _ptr(rho), int(bool(adv_den)),
                                              int(bool(horz_only)), int(bool(use_dqdt)), _ptr(out)))
        return out

    def _out(self, out, name):
        """A mass-grid output: given (device fields: required), or a new host array."""
        if self.device_fields:
            if out is None:
                raise ValueError(f"{name}: pass `out` (a device array) when fields live on the device")
            return self._arr(out, self.mass_shape, "out")
        if out is None:
            return np.zeros(self.mass_shape, dtype=self.dtype)
        if not (isinstance(out, np.ndarray) and out.dtype == self.dtype and out.flags.c_contiguous
                and out.shape == self.mass_shape):
            raise ValueError(f"{name}: need a C-contiguous {self.dtype} array of shape {self.mass_shape}")
        return out

    def balance_uvw(self, u, v, w, density=None, adv_den=True, use_dqdt=False):
        """balance_uvw(domain, adv_den, update_in) (wind.F90:123-191): w (updated in
        place) from the horizontal divergence of u, v by calc_w (:193-234)."""
        uu = self._arr(u, self.u_shape, "u")
        vv = self._arr(v, self.v_shape, "v")
        ww = self._out(w, "balance_uvw")
        rho = self._arr(density, self.mass_shape, "density")
        nat.check(self.lib.windhip_balance_uvw(self.h, _ptr(uu), _ptr(vv), _ptr(ww), _ptr(rho),
                                               int(bool(adv_den)), int(bool(use_dqdt))))
        return w if self.device_fields else ww

    def calc_w_real(self, u, v, w_grid=None, out=None):
        """calc_w_real (wind.F90:1175-1227): the real vertical velocity on the mass
        grid; w_grid None is zero."""
        uu = self._arr(u, self.u_shape, "u")
        vv = self._arr(v, self.v_shape, "v")
        wg = self._arr(w_grid, self.mass_shape, "w_grid")
        o = self._out(out, "calc_w_real")
        nat.check(self.lib.windhip_calc_w_real(self.h, _ptr(uu), _ptr(vv), _ptr(wg), _ptr(o)))
        return out if self.device_fields else o

    def idealized_wgrid(self, u, v, w_real, out=None):
        """calc_idealized_wgrid (wind.F90:940-998): the grid-relative w that
        reproduces w_real for the given u, v."""
        uu = self._arr(u, self.u_shape, "u")
        vv = self._arr(v, self.v_shape, "v")
        wr = self._arr(w_real, self.mass_shape, "w_real")
        o = self._out(out, "idealized_wgrid")
        nat.check(self.lib.windhip_idealized_wgrid(self.h, _ptr(uu), _ptr(vv), _ptr(wr), _ptr(o)))
        return out if self.device_fields else o

    def get_lambda(self) -> np.ndarray:
        out = np.zeros(self.rows_shape)
        nat.check(self.lib.windhip_get_lambda(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def set_lambda(self, lam) -> None:
        lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(self.rows_shape)
        nat.check(self.lib.windhip_set_lambda(self.h, lam.ctypes.data_as(C.POINTER(C.c_double))))

    def set_multigrid(self, pre: int = 2, post: int = 2, coarse_pre: int = 1, coarse_post: int = 1, depth: int = 0,
                      omega: float = 0.9) -> None:
        """V-cycle of the MIXED_MG inner preconditioner (no counterpart in the reference)."""
        nat.check(self.lib.windhip_set_multigrid(self.h, int(pre), int(post), int(coarse_pre), int(coarse_post),
                                                 int(depth), float(omega)))

    def set_mg_fusion(self, fused_transfers: int = 1, onepass: bool = True) -> None:
        """MIXED_MG kernel forms. fused_transfers: 1 (or True, default), the
        one-launch descent/return between the wavefront-per-column levels; 2, as
        1 with the cycle from the first such level down as one persistent launch
        (grid-wide barriers between its steps); 0 (or False), separate kernels.
        onepass: the one-pass smoothing steps (on by default). No counterpart in
        the reference."""
        nat.check(self.lib.windhip_set_mg_fusion(self.h, int(fused_transfers), int(bool(onepass))))

    def set_mg_galerkin(self, first_level: int = 2) -> None:
        """Coarse levels from `first_level` on take the Galerkin product R A P of the
        finer level's operator (0: every level rediscretised); before the first
        MIXED_MG solve. No counterpart in the reference."""
        nat.check(self.lib.windhip_set_mg_galerkin(self.h, int(first_level)))

    def mg_level(self, level: int):
        """Coarse level `level` of the last MIXED_MG solve: ((nx, ny, mz), coefficients [6, mz, ny+2, nx+2]
        = B, C, H, I, L, M of the rediscretised operator; None on a Galerkin level)."""
        dims = (C.c_int * 3)()
        nat.check(self.lib.windhip_mg_level(self.h, int(level), dims, None))
        nx, ny, mz = dims[0], dims[1], dims[2]
        if self.lib.windhip_mg_stencil(self.h, int(level), None) == 0:
            return (nx, ny, mz), None
        out = np.zeros((6, mz, ny + 2, nx + 2), dtype=np.float32)
        nat.check(self.lib.windhip_mg_level(self.h, int(level), dims, out.ctypes.data_as(C.POINTER(C.c_float))))
        return (nx, ny, mz), out

    def mg_stencil(self, level: int):
        """The 27 coefficients [27, mz, ny+2, nx+2] of Galerkin level `level`
        (coefficient 9 (dk+1) + 3 (dj+1) + (di+1)); None on a rediscretised level."""
        dims = (C.c_int * 3)()
        nat.check(self.lib.windhip_mg_level(self.h, int(level), dims, None))
        if self.lib.windhip_mg_stencil(self.h, int(level), None) != 0:
            return None
        nx, ny, mz = dims[0], dims[1], dims[2]
        out = np.zeros((27, mz, ny + 2, nx + 2), dtype=np.float32)
        nat.check(self.lib.windhip_mg_stencil(self.h, int(level), out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def mg_smooth(self, level: int, fused, b, y=None):
        """One damped line-smoothing step on coarse level `level` of the last MIXED_MG
        solve: y + omega L^-1 (b - A y), FP32 [mz, ny+2, nx+2] (no counterpart in
        the reference). fused: True or 1, the one-launch form; False or 0, the
        residual + line-solve kernels on FP32 rows; 2, those kernels on the
        level's bf16 rows (levels of up to 44 rows per column); 3 and 4, the
        one-pass step (residual and line solve in one kernel) on the bf16 (3)
        or FP32 (4) rows, for y given."""
        dims = (C.c_int * 3)()
        nat.check(self.lib.windhip_mg_level(self.h, int(level), dims, None))
        shape = (dims[2], dims[1] + 2, dims[0] + 2)
        fpp = C.POINTER(C.c_float)
        bb = np.ascontiguousarray(b, dtype=np.float32).reshape(shape)
        yy = None if y is None else np.ascontiguousarray(y, dtype=np.float32).reshape(shape)
        out = np.zeros(shape, dtype=np.float32)
        nat.check(self.lib.windhip_mg_smooth(self.h, int(level), int(fused), bb.ctypes.data_as(fpp),
                                             None if yy is None else yy.ctypes.data_as(fpp),
                                             out.ctypes.data_as(fpp)))
        return out

    def mg_cycle(self, level: int, b, fused: int = 1):
        """One V-cycle of the last MIXED_MG solve's hierarchy from coarse level
        `level` for the rhs b, FP32 [mz, ny+2, nx+2] (no counterpart in the
        reference); fused 1 (or True) selects the one-launch descent and return
        on the levels of one wavefront per column, 2 in addition the persistent
        launch from the first such level down, 0 (or False) separate kernels."""
        dims = (C.c_int * 3)()
        nat.check(self.lib.windhip_mg_level(self.h, int(level), dims, None))
        shape = (dims[2], dims[1] + 2, dims[0] + 2)
        fpp = C.POINTER(C.c_float)
        bb = np.ascontiguousarray(b, dtype=np.float32).reshape(shape)
        out = np.zeros(shape, dtype=np.float32)
        nat.check(self.lib.windhip_mg_cycle(self.h, int(level), int(fused), bb.ctypes.data_as(fpp),
                                            out.ctypes.data_as(fpp)))
        return out

    def probe_operator(self, alpha, density=None, adv_den=True) -> float:
        """calibrate_projection_operator (wind.F90:1010-1099): A = 2 D(G(.))."""
        a = self._arr(alpha, self.mass_shape, "alpha")
        rho = self._arr(density, self.mass_shape, "density")
        leak = C.c_double()
        nat.check(self.lib.windhip_probe_operator(self.h, _ptr(a), _ptr(rho), int(bool(adv_den)), C.byref(leak)))
        return leak.value

    # ------------------------------------------------------------ diagnostics
    ROWS = {"rhs": 0, "dinv": 1, "B": 2, "C": 3, "H": 4, "I": 5, "L": 6, "M": 7, "lambda": 8}

    def set_alpha(self, alpha) -> None:
        a = self._arr(alpha, self.mass_shape, "alpha")
        nat.check(self.lib.windhip_set_alpha(self.h, _ptr(a)))

    def apply_operator(self, x, precision: int = nat.PLAIN_F64) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.rows_shape)
        y = np.zeros(self.rows_shape)
        dp = C.POINTER(C.c_double)
        nat.check(self.lib.windhip_apply_operator(self.h, precision, x.ctypes.data_as(dp), y.ctypes.data_as(dp)))
        return y

    def get_rows(self, which: str) -> np.ndarray:
        out = np.zeros(self.rows_shape)
        nat.check(self.lib.windhip_get_rows(self.h, self.ROWS[which], out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def profile(self, reset: bool = True) -> dict:
        """Per-kernel (calls, total ms) of the solves run with profile=1."""
        calls = (C.c_double * nat.PROFILE_SLOTS)()
        ms = (C.c_double * nat.PROFILE_SLOTS)()
        nat.check(self.lib.windhip_profile(self.h, int(reset), calls, ms))
        return {name: (calls[k], ms[k]) for k, name in nat.PROFILE_KINDS.items() if calls[k] > 0}

    def time_kernel(self, which: int = 0, reps: int = 20) -> float:
        ms = C.c_double()
        nat.check(self.lib.windhip_time_kernel(self.h, which, reps, C.byref(ms)))
        return ms.value

    def memory(self):
        a, b = C.c_double(), C.c_double()
        nat.check(self.lib.windhip_memory(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.windhip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------- module-style interface
@dataclass
class Domain:
    """The parts of HICAR's domain_t the wind solve reads: the tile size, dx,
    nest index and the metric fields (numpy, layout as above)."""
    nx: int
    ny: int
    nz: int
    dx: float
    metrics: dict
    nest_indx: int = 1


MAX_NESTS = 4  # wind_iterative.F90:139
_cache: dict = {}
_active = {"nest": -1}


def init_iter_winds(domain: Domain, dtype=np.float64) -> WindSolver:
    """init_iter_winds(domain, options) (wind_iterative.F90:215-291): one solver
    state per nest, switched in O(1) when the nest changes."""
    n = domain.nest_indx
    if n in _cache:
        _active["nest"] = n
        return _cache[n]
    if len(_cache) >= MAX_NESTS:
        raise ValueError(f"at most {MAX_NESTS} nests")
    s = WindSolver(domain.nx, domain.ny, domain.nz, domain.dx, dtype)
    s.set_metrics(domain.metrics)
    _cache[n] = s
    _active["nest"] = n
    return s


def calc_iter_winds(domain: Domain, alpha, div, u, v, w, density=None, adv_den=True, **opts) -> dict:
    """calc_iter_winds(domain, alpha_in, div_in, adv_den) (wind_iterative.F90:295)."""
    s = _cache.get(domain.nest_indx)
    if s is None:
        raise RuntimeError("calc_iter_winds: init_iter_winds has not been called for this nest")
    return s.calc_iter_winds(alpha, div, u, v, w, density, adv_den, **opts)


def finalize_iter_winds() -> None:
    """finalize_iter_winds (wind_iterative.F90:2142-2267): release every nest."""
    for s in _cache.values():
        s.close()
    _cache.clear()
    _active["nest"] = -1
