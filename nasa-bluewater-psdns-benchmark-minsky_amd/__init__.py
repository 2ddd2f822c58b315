"""MI355X (gfx950) implementation of the PSDNS RK2 pseudo-spectral step.

The product is the C-ABI library lib/libpsdns_gpu.so (include/psdns_gpu.h)
and the Fortran ISO_C_BINDING module + driver (fortran/).  This Python module
is a thin ctypes binding to the same C ABI, used by the tests and bench.py;
it never computes anything itself, and it raises if the library is missing.

Host arrays follow the reference layouts (C order, fastest index last):
    cylinder uy(ny, num_al, nv) -> numpy (nv, num_al, n) complex
    physical ux(nx, nz, ny, nv) -> numpy (nv, n, n, n)   real, [v][y][z][x]
"""
import ctypes
import pathlib

import numpy as np

import os

PKG = pathlib.Path(__file__).resolve().parent
# PSDNS_GPU_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = pathlib.Path(os.environ.get("PSDNS_GPU_LIB", PKG / "lib" / "libpsdns_gpu.so"))
DRIVER_PATH = PKG / "bin" / "psdns_driver"

FIELD_U, FIELD_UN, FIELD_UX = 0, 1, 2

# exported C-ABI symbols (include/psdns_gpu.h)
EXPORTS = (
    "psdns_create", "psdns_destroy", "psdns_geometry", "psdns_columns", "psdns_put_state",
    "psdns_get_state", "psdns_set_shifts", "psdns_stage_nonlinear", "psdns_stage_update",
    "psdns_ek_shell", "psdns_ek_force", "psdns_shell_stats", "psdns_shell_stats_noalloc",
    "psdns_init_field", "psdns_put_field",
    "psdns_get_field", "psdns_phase1", "psdns_itransform_vel", "psdns_realspace_vel", "psdns_transform",
    "psdns_wavespace_vel", "psdns_proc4a", "psdns_proc4b", "psdns_kxtran", "psdns_xktran",
    "psdns_stage_times", "psdns_vg_moments", "psdns_outpen", "psdns_inpen", "psdns_rk4_update",
    "psdns_set_epfor", "psdns_epfor_tally", "psdns_rk4_forcing", "psdns_pass_times", "psdns_lag_put",
    "psdns_lag_get", "psdns_lag_spline", "psdns_lag_interp", "psdns_lag_advpos", "psdns_last_error",
)
# learned SGS closure (include/psdns_sgs.h)
SGS_EXPORTS = (
    "psdns_sgs_param_count", "psdns_sgs_create", "psdns_sgs_destroy", "psdns_sgs_set_params",
    "psdns_sgs_get_params", "psdns_sgs_filter", "psdns_sgs_features", "psdns_sgs_put_fields",
    "psdns_sgs_get_field", "psdns_sgs_add_snapshot", "psdns_sgs_snapshots", "psdns_sgs_clear_snapshots",
    "psdns_sgs_target_scale", "psdns_sgs_gradient", "psdns_sgs_train", "psdns_sgs_predict", "psdns_sgs_eval",
    "psdns_sgs_add_stress", "psdns_sgs_features_stage", "psdns_sgs_attach_stress",
)
SGS_LAMBDA, SGS_BASIS, SGS_TARGET, SGS_PREDICTED, SGS_GRADIENT = range(5)


class PsdnsError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("nc", ctypes.c_int),
        ("prec", ctypes.c_int), ("kshift", ctypes.c_int), ("mask_kind", ctypes.c_int),
        ("viscos", ctypes.c_double), ("pr", ctypes.c_double * 8), ("grad", (ctypes.c_double * 3) * 8),
    ]


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


def build(verbose=False):
    """Compile lib/libpsdns_gpu.so and the Fortran driver (build.py)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("_psdns_build", PKG / "build.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.build(verbose)


def load_library():
    """Load lib/libpsdns_gpu.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise PsdnsError(f"{LIB_PATH} is missing: run build() (hipcc --offload-arch=gfx950) first")
    L = ctypes.CDLL(str(LIB_PATH))
    i, d = ctypes.c_int, ctypes.c_double
    sig = {
        "psdns_create": [ctypes.POINTER(Params), ctypes.POINTER(_vp)],
        "psdns_destroy": [_vp],
        "psdns_geometry": [_vp, _ip, _ip],
        "psdns_columns": [_vp, _ip, _ip, _ip],
        "psdns_put_state": [_vp, _vp, i],
        "psdns_get_state": [_vp, _vp, i],
        "psdns_set_shifts": [_vp, _dp],
        "psdns_stage_nonlinear": [_vp, i, _dp],
        "psdns_stage_update": [_vp, i, d, _dp, _dp, _dp, _vp],
        "psdns_ek_shell": [_vp, i, _dp],
        "psdns_ek_force": [_vp, i, i, _dp, _dp],
        "psdns_shell_stats": [_vp, _dp, _dp, i, _dp, _dp],
        "psdns_shell_stats_noalloc": [_vp, _dp, _dp, i, _dp, _dp],
        "psdns_vg_moments": [_vp, _dp, _dp],
        "psdns_init_field": [_vp, ctypes.c_ulonglong, d, d],
        "psdns_put_field": [_vp, i, i, _vp],
        "psdns_get_field": [_vp, i, i, _vp],
        "psdns_phase1": [_vp, i],
        "psdns_itransform_vel": [_vp],
        "psdns_realspace_vel": [_vp, i, _dp],
        "psdns_transform": [_vp],
        "psdns_wavespace_vel": [_vp, i, d],
        "psdns_proc4a": [_vp, _dp, _dp, _dp],
        "psdns_proc4b": [_vp],
        "psdns_kxtran": [_vp, i],
        "psdns_xktran": [_vp, i],
        "psdns_stage_times": [_vp, i, _dp],
        "psdns_outpen": [_vp, ctypes.c_char_p, i],
        "psdns_inpen": [_vp, ctypes.c_char_p],
        "psdns_rk4_update": [_vp, i, d, _dp, _dp, _dp],
        "psdns_set_epfor": [_vp, i, d],
        "psdns_epfor_tally": [_vp, _dp],
        "psdns_rk4_forcing": [_vp, _vp],
        "psdns_pass_times": [_vp, i, _dp, i],
        "psdns_lag_put": [_vp, _dp, i, i],
        "psdns_lag_get": [_vp, _dp, i],
        "psdns_lag_spline": [_vp],
        "psdns_lag_interp": [_vp],
        "psdns_lag_advpos": [_vp, i, d, d, d],
        "psdns_sgs_param_count": [],
        "psdns_sgs_create": [i, i, ctypes.POINTER(_vp)],
        "psdns_sgs_destroy": [_vp],
        "psdns_sgs_set_params": [_vp, _fp, i],
        "psdns_sgs_get_params": [_vp, _fp, i],
        "psdns_sgs_filter": [_vp, _vp, _vp],
        "psdns_sgs_features": [_vp, _vp],
        "psdns_sgs_put_fields": [_vp, _fp, _fp, _fp],
        "psdns_sgs_get_field": [_vp, i, _fp],
        "psdns_sgs_add_snapshot": [_vp],
        "psdns_sgs_snapshots": [_vp, _ip],
        "psdns_sgs_clear_snapshots": [_vp],
        "psdns_sgs_target_scale": [_vp, _dp],
        "psdns_sgs_gradient": [_vp, i, ctypes.c_ulonglong, ctypes.c_longlong, d, _fp, _dp],
        "psdns_sgs_train": [_vp, i, i, d, ctypes.c_ulonglong, d, _dp],
        "psdns_sgs_predict": [_vp, i],
        "psdns_sgs_eval": [_vp, d, _dp],
        "psdns_sgs_add_stress": [_vp, _vp],
        "psdns_sgs_features_stage": [_vp, _vp, i],
        "psdns_sgs_attach_stress": [_vp, _vp, i],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.psdns_last_error.restype = ctypes.c_char_p
    L.psdns_last_error.argtypes = []
    _lib = L
    return L


def _dptr(a):
    return a.ctypes.data_as(_dp)


class Solver:
    """One device context (psdns_ctx) for an n^3 box at precision prec (4|8)."""

    def __init__(self, n, prec=8, viscos=0.01, nc=0, pr=None, grad=None):
        """nc passive scalars with Prandtl/Schmidt numbers pr[nc] and mean
        gradients grad[nc][3] (state slots 3 .. 2 + nc)."""
        self.L = load_library()
        self.n, self.prec, self.viscos, self.nc = n, prec, viscos, nc
        p = Params(nx=n, ny=n, nz=n, nc=nc, prec=prec, kshift=1, mask_kind=0, viscos=viscos)
        for i in range(nc):
            p.pr[i] = (pr or [1.0] * nc)[i]
            for d in range(3):
                p.grad[i][d] = (grad or [(0.0, 0.0, 0.0)] * nc)[i][d]
        h = _vp()
        self._chk(self.L.psdns_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        na, mx = ctypes.c_int(), ctypes.c_int()
        self._chk(self.L.psdns_geometry(self.h, ctypes.byref(na), ctypes.byref(mx)))
        self.num_al, self.max_al_x = na.value, mx.value
        self.cdtype = np.complex128 if prec == 8 else np.complex64
        self.rdtype = np.float64 if prec == 8 else np.float32

    def _chk(self, rc):
        if rc != 0:
            raise PsdnsError(f"psdns error {rc}: {self.L.psdns_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.psdns_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def columns(self):
        cx = np.zeros(self.num_al, np.int32)
        cz = np.zeros(self.num_al, np.int32)
        cut = np.zeros(self.max_al_x, np.int32)
        self._chk(self.L.psdns_columns(self.h, cx.ctypes.data_as(_ip), cz.ctypes.data_as(_ip), cut.ctypes.data_as(_ip)))
        return cx, cz, cut

    # state and fields
    def put_state(self, uny):
        """velocity (3) and, with scalars, the nc scalar fields"""
        a = np.ascontiguousarray(uny[:3 + self.nc], self.cdtype)
        if a.shape[0] != 3 + self.nc:
            raise PsdnsError(f"state needs {3 + self.nc} variables, got {a.shape[0]}")
        self._chk(self.L.psdns_put_state(self.h, a.ctypes.data, 0))

    def get_state(self):
        a = np.zeros((3 + self.nc, self.num_al, self.n), self.cdtype)
        self._chk(self.L.psdns_get_state(self.h, a.ctypes.data, 0))
        return a

    def outpen(self, path, file_prec=8):
        """Write un as the reference's pencil checkpoint file
        (homogeneous/outpen_write.f:59-76), complex*16 (file_prec 8) or
        complex*8 (4), big-endian Fortran records."""
        self._chk(self.L.psdns_outpen(self.h, os.fsencode(path), file_prec))

    def inpen(self, path):
        """Read such a file (either precision) into un, and u = un
        (homogeneous/inpen_read.f)."""
        self._chk(self.L.psdns_inpen(self.h, os.fsencode(path)))

    def put_field(self, which, arr):
        dt = self.rdtype if which == FIELD_UX else self.cdtype
        a = np.ascontiguousarray(arr, dt)
        self._chk(self.L.psdns_put_field(self.h, which, a.shape[0], a.ctypes.data))

    def get_field(self, which, nv):
        if which == FIELD_UX:
            a = np.zeros((nv, self.n, self.n, self.n), self.rdtype)
        else:
            a = np.zeros((nv, self.num_al, self.n), self.cdtype)
        self._chk(self.L.psdns_get_field(self.h, which, nv, a.ctypes.data))
        return a

    def set_shifts(self, gsh):
        g = np.ascontiguousarray(np.asarray(gsh, np.float64).reshape(6))
        self._chk(self.L.psdns_set_shifts(self.h, _dptr(g)))

    def init_field(self, seed=122929298, kpeak=3.0, tke=0.5):
        self._chk(self.L.psdns_init_field(self.h, seed, kpeak, tke))

    # fused stage entry points
    def stage_nonlinear(self, kstep):
        vm = np.zeros(1)
        self._chk(self.L.psdns_stage_nonlinear(self.h, kstep, _dptr(vm)))
        return float(vm[0])

    def stage_update(self, kstep, dt, xfac=None, yfac=None, zfac=None, forcing=None):
        """forcing: EPFOR increments for(kfor, k2fo, k2fo, 3) as a (3, k2fo,
        k2fo, kfor) complex array (needs set_epfor)"""
        f = None if forcing is None else self._forcing(forcing)
        fp = None if f is None else f.ctypes.data
        if xfac is None:
            self._chk(self.L.psdns_stage_update(self.h, kstep, dt, None, None, None, fp))
        else:
            xf, yf, zf = (np.ascontiguousarray(f, np.float64) for f in (xfac, yfac, zfac))
            self._chk(self.L.psdns_stage_update(self.h, kstep, dt, _dptr(xf), _dptr(yf), _dptr(zf), fp))

    def _forcing(self, forcing):
        f = np.ascontiguousarray(forcing, np.complex128)
        k = getattr(self, "epf_kfor", 0)
        if k == 0 or f.shape != (3, 2 * k - 1, 2 * k - 1, k):
            raise PsdnsError(f"forcing needs set_epfor and shape (3, 2kfor-1, 2kfor-1, kfor), got {f.shape}")
        return f

    def set_epfor(self, kfor, kforce):
        """EPFOR forcing box kfor and radius kforce (0 turns it off)"""
        self._chk(self.L.psdns_set_epfor(self.h, kfor, kforce))
        self.epf_kfor = kfor

    def epfor_tally(self):
        """efkz(nxh, 3) of the last proc4a / advanc_1 as a (3, n/2) array"""
        e = np.zeros((3, self.n // 2))
        self._chk(self.L.psdns_epfor_tally(self.h, _dptr(e)))
        return e

    def rk4_forcing(self, forcing):
        """EPFOR increments for the following rk4_update calls (None clears)"""
        f = None if forcing is None else self._forcing(forcing)
        self._chk(self.L.psdns_rk4_forcing(self.h, None if f is None else f.ctypes.data))

    def rk4_update(self, kstep, dt, xfac=None, yfac=None, zfac=None):
        """advanc_kstep (homogeneous/advanc_1.f .. advanc_4.f) after
        stage_nonlinear(kstep); factors on kstep 1 only."""
        args = [None if a is None else _dptr(np.ascontiguousarray(a, np.float64)) for a in (xfac, yfac, zfac)]
        self._chk(self.L.psdns_rk4_update(self.h, kstep, dt, *args))

    def step_factors(self, dt):
        """step.f:366-385 integrating factors (host side, as the kept driver)."""
        n = self.n
        k = np.arange(n)
        kw = np.where(k > n // 2, k - n, k).astype(np.float64)
        kx = np.arange(n // 2, dtype=np.float64)
        return (np.exp(-self.viscos * kx * kx * dt), np.exp(-self.viscos * kw * kw * dt),
                np.exp(-self.viscos * kw * kw * dt))

    def ek_shell(self, kfor):
        ek = np.zeros(kfor)
        self._chk(self.L.psdns_ek_shell(self.h, kfor, _dptr(ek)))
        return ek

    def ek_force(self, kf_shell, kfor, ekinf):
        e = np.ascontiguousarray(ekinf, np.float64)
        out = np.zeros(kfor)
        self._chk(self.L.psdns_ek_force(self.h, kf_shell, kfor, _dptr(e), _dptr(out)))
        return out

    def shell_stats(self, nshell=None, noalloc=False):
        """FP64 E(k), D(k), TKE, epsilon; noalloc selects the entry point
        that allocates nothing (psdns_shell_stats_noalloc)"""
        nshell = self.n if nshell is None else nshell
        ek, dk = np.zeros(max(nshell, 1)), np.zeros(max(nshell, 1))
        t, e = np.zeros(1), np.zeros(1)
        f = self.L.psdns_shell_stats_noalloc if noalloc else self.L.psdns_shell_stats
        self._chk(f(self.h, _dptr(ek), _dptr(dk), nshell, _dptr(t), _dptr(e)))
        return ek[:nshell], dk[:nshell], float(t[0]), float(e[0])

    def vg_moments(self):
        """Moments 1..8 of du/dx and dw/dz over the grid (vgmomt.f:68-143)."""
        du, dw = np.zeros(8), np.zeros(8)
        self._chk(self.L.psdns_vg_moments(self.h, _dptr(du), _dptr(dw)))
        return du, dw

    def stage_times(self, enable=True):
        t = np.zeros(6)
        self._chk(self.L.psdns_stage_times(self.h, int(enable), _dptr(t)))
        return t

    PASS_NAMES = ("y_inverse", "z_inverse", "x_pass", "z_forward", "y_forward_update", "sc_grad_y",
                  "sc_z_inverse", "sc_x_pass", "sc_z_forward", "sc_y_forward", "sc_update")

    def pass_times(self, enable=True):
        """device ms per RK substep (4 rows) and pass (PASS_NAMES), cleared on read"""
        t = np.zeros(44)
        self._chk(self.L.psdns_pass_times(self.h, int(enable), _dptr(t), 44))
        return t.reshape(4, 11)

    # Lagrangian particles: pp is the reference's pp(np, 9) as a (9, np)
    # float64 array (positions in grid units, velocities, saved positions)
    def lag_put(self, pp):
        a = np.ascontiguousarray(pp, np.float64)
        self._chk(self.L.psdns_lag_put(self.h, _dptr(a), a.shape[1], a.shape[0]))
        self.lag_np = a.shape[1]

    def lag_get(self):
        a = np.zeros((9, self.lag_np))
        self._chk(self.L.psdns_lag_get(self.h, _dptr(a), self.lag_np))
        return a

    def lag_spline(self):
        self._chk(self.L.psdns_lag_spline(self.h))

    def lag_interp(self):
        self._chk(self.L.psdns_lag_interp(self.h))

    def lag_advpos(self, pstep, dt, cof1, cof2):
        self._chk(self.L.psdns_lag_advpos(self.h, pstep, dt, cof1, cof2))

    def lag_step(self, dt_prev):
        """partsp.f's RK2 particle step after set_shifts of the new step:
        interpolation only on the first call (dt_prev None), else
        rk_advpos(1, 1/2, 1/2), spline + interpolation, rk_advpos(2, 0, 1/2)"""
        if dt_prev is None:
            self.lag_spline()
            self.lag_interp()
            return
        self.lag_advpos(1, dt_prev, 0.5, 0.5)
        self.lag_spline()
        self.lag_interp()
        self.lag_advpos(2, dt_prev, 0.0, 0.5)

    # per-routine entry points
    def phase1(self, kstep):
        self._chk(self.L.psdns_phase1(self.h, kstep))

    def itransform_vel(self):
        self._chk(self.L.psdns_itransform_vel(self.h))

    def realspace_vel(self, rkstep):
        vm = np.zeros(1)
        self._chk(self.L.psdns_realspace_vel(self.h, rkstep, _dptr(vm)))
        return float(vm[0])

    def transform(self):
        self._chk(self.L.psdns_transform(self.h))

    def wavespace_vel(self, kstep, dt2):
        self._chk(self.L.psdns_wavespace_vel(self.h, kstep, dt2))

    def proc4a(self, xfac, yfac, zfac):
        xf, yf, zf = (np.ascontiguousarray(f, np.float64) for f in (xfac, yfac, zfac))
        self._chk(self.L.psdns_proc4a(self.h, _dptr(xf), _dptr(yf), _dptr(zf)))

    def proc4b(self):
        self._chk(self.L.psdns_proc4b(self.h))

    def kxtran(self, nv):
        self._chk(self.L.psdns_kxtran(self.h, nv))

    def xktran(self, nv):
        self._chk(self.L.psdns_xktran(self.h, nv))


class Run4:
    """RK4 time loop (rkmethod = 4, timestep_module.f:293-310): four
    substeps, dt and the integrating factors from substep 1, FEK forcing
    after the step.  Shift offsets gsh(3,2) from the caller serve stages
    1, 2 and again 3, 4 (shifts.f:133-137).  Passive scalars of the solver
    step with the velocity."""

    def __init__(self, solver, viscos, cfl, kf_shell=0, ekinf=None, epfor=None):
        self.s, self.viscos, self.cfl = solver, viscos, cfl
        self.kf_shell = kf_shell
        self.kfor = kf_shell + 1 if kf_shell > 0 else 0
        self.ekinf = ekinf if ekinf is not None or not self.kfor else solver.ek_shell(self.kfor)
        self.dts = []
        # EPFOR: (kfor, kforce, forcing) with forcing(istep) the increments of step istep
        self.epfor = epfor
        if epfor is not None:
            solver.set_epfor(epfor[0], epfor[1])

    def step(self, gsh):
        s = self.s
        s.set_shifts(gsh)
        if self.epfor is not None:
            s.rk4_forcing(self.epfor[2](len(self.dts)))
        for k in range(1, 5):
            vm = s.stage_nonlinear(k)
            if k == 1:
                self.velmax = vm
                self.dt = self.cfl / vm
                xf, yf, zf = s.step_factors(self.dt)
                s.rk4_update(1, self.dt, xf, yf, zf)
            else:
                s.rk4_update(k, self.dt)
        if self.kfor:
            s.ek_force(self.kf_shell, self.kfor, self.ekinf)
        self.dts.append(self.dt)


class Run:
    """RK2 time loop over the fused entry points, as the kept driver runs it
    (main.f:392-685 / timestep_module.f:25-481): shifts, then per stage
    stage_nonlinear -> (stage 1: dt = cfl/velmax, integrating factors) ->
    stage_update, then FEK forcing.  Shift offsets come from the caller."""

    def __init__(self, solver, viscos, cfl, kf_shell=0, ekinf=None, epfor=None, particles=None):
        self.s, self.viscos, self.cfl = solver, viscos, cfl
        self.kf_shell = kf_shell
        self.kfor = kf_shell + 1 if kf_shell > 0 else 0
        self.ekinf = ekinf if ekinf is not None or not self.kfor else solver.ek_shell(self.kfor)
        self.dts = []
        # EPFOR: (kfor, kforce, forcing) with forcing(istep) the increments
        # for(kfor, k2fo, k2fo, 3) of step istep (force.f, called after step)
        self.epfor = epfor
        if epfor is not None:
            solver.set_epfor(epfor[0], epfor[1])
        # Lagrangian particles pp (9, np): stepped at the start of each step
        # (lagpart from realspace at kstep 1, with the previous dt)
        self.particles = particles is not None
        if self.particles:
            solver.lag_put(particles)

    def step(self, gsh):
        s = self.s
        s.set_shifts(gsh)
        if self.particles:
            s.lag_step(self.dts[-1] if self.dts else None)
        vm = s.stage_nonlinear(1)
        self.velmax = vm
        self.dt = self.cfl / vm
        xf, yf, zf = s.step_factors(self.dt)
        f = None if self.epfor is None else self.epfor[2](len(self.dts))
        s.stage_update(1, self.dt, xf, yf, zf, forcing=f)
        s.stage_nonlinear(2)
        s.stage_update(2, self.dt, xf, yf, zf, forcing=f)
        if self.kfor:
            s.ek_force(self.kf_shell, self.kfor, self.ekinf)
        self.dts.append(self.dt)


def _fptr(a):
    return a.ctypes.data_as(_fp)


class Closure:
    """Learned SGS closure for an LES grid n^3 (include/psdns_sgs.h): the
    network parameters and Adam state, the features/target/prediction of
    the current field and a store of training snapshots, all on the GPU."""

    def __init__(self, n, max_snapshots=0):
        self.L = load_library()
        self.n = n
        self.P = n ** 3
        self.nparam = self.L.psdns_sgs_param_count()
        h = _vp()
        self._chk(self.L.psdns_sgs_create(n, max_snapshots, ctypes.byref(h)))
        self.h = h

    def _chk(self, rc):
        if rc != 0:
            raise PsdnsError(f"psdns error {rc}: {self.L.psdns_last_error().decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.psdns_sgs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, theta):
        a = np.ascontiguousarray(theta, np.float32)
        self._chk(self.L.psdns_sgs_set_params(self.h, _fptr(a), a.size))

    def get_params(self):
        a = np.zeros(self.nparam, np.float32)
        self._chk(self.L.psdns_sgs_get_params(self.h, _fptr(a), a.size))
        return a

    def filter(self, fine, coarse):
        """box filter of fine's ux slots 0..2 into coarse's ux slots 0..2 + the target"""
        self._chk(self.L.psdns_sgs_filter(self.h, fine.h, coarse.h))

    def features(self, solver, kstep=0):
        """features of solver's state u: as it is (kstep 0) or on stage
        kstep's shifted grid (1|2)"""
        if kstep:
            self._chk(self.L.psdns_sgs_features_stage(self.h, solver.h, kstep))
        else:
            self._chk(self.L.psdns_sgs_features(self.h, solver.h))

    def attach(self, solver, kstep):
        """the predicted stress goes into the next stage_nonlinear(kstep)"""
        self._chk(self.L.psdns_sgs_attach_stress(self.h, solver.h, kstep))

    def put_fields(self, lam, basis, tau=None):
        la = np.ascontiguousarray(lam, np.float32)
        ba = np.ascontiguousarray(basis, np.float32)
        ta = None if tau is None else np.ascontiguousarray(tau, np.float32)
        self._chk(self.L.psdns_sgs_put_fields(self.h, _fptr(la), _fptr(ba), None if ta is None else _fptr(ta)))

    _SHAPES = {SGS_LAMBDA: (5,), SGS_BASIS: (4, 6), SGS_TARGET: (6,), SGS_PREDICTED: (6,), SGS_GRADIENT: (3, 3)}

    def field(self, which):
        n = self.n
        a = np.zeros(self._SHAPES[which] + (n, n, n), np.float32)
        self._chk(self.L.psdns_sgs_get_field(self.h, which, _fptr(a)))
        return a

    def add_snapshot(self):
        self._chk(self.L.psdns_sgs_add_snapshot(self.h))

    def snapshots(self):
        c = ctypes.c_int()
        self._chk(self.L.psdns_sgs_snapshots(self.h, ctypes.byref(c)))
        return c.value

    def clear_snapshots(self):
        self._chk(self.L.psdns_sgs_clear_snapshots(self.h))

    def target_scale(self):
        s = np.zeros(1)
        self._chk(self.L.psdns_sgs_target_scale(self.h, _dptr(s)))
        return float(s[0])

    def gradient(self, batch, seed, step, sigma2):
        g = np.zeros(self.nparam, np.float32)
        loss = np.zeros(1)
        self._chk(self.L.psdns_sgs_gradient(self.h, batch, seed, step, sigma2, _fptr(g), _dptr(loss)))
        return g, float(loss[0])

    def train(self, nsteps, batch, lr, seed, sigma2):
        losses = np.zeros(max(nsteps, 1))
        self._chk(self.L.psdns_sgs_train(self.h, nsteps, batch, lr, seed, sigma2, _dptr(losses)))
        return losses[:nsteps]

    def predict(self, bf16=False):
        self._chk(self.L.psdns_sgs_predict(self.h, int(bool(bf16))))

    def eval(self, sigma2):
        loss = np.zeros(1)
        self._chk(self.L.psdns_sgs_eval(self.h, sigma2, _dptr(loss)))
        return float(loss[0])

    def add_stress(self, solver):
        self._chk(self.L.psdns_sgs_add_stress(self.h, solver.h))


def training_snapshot(closure, fine, coarse):
    """One training snapshot from the fine DNS state: its physical velocity
    (kxtran), box-filtered onto the coarse grid (velocity + target stress),
    the coarse spectral state (xktran) and its features."""
    fine.kxtran(3)
    closure.filter(fine, coarse)
    coarse.xktran(3)
    closure.features(coarse)
    closure.add_snapshot()


class Les:
    """RK2 time loop of a coarse-grid LES with the closure's stress added to
    the products; closure None runs the same grid without a model.
    fused: the fused entry points (stage_nonlinear / stage_update, the
    stress added in the x pass, features on the stage's shifted grid);
    otherwise the per-routine step (phase1, itransform_vel, realspace_vel,
    transform, wavespace_vel, proc4a/proc4b per stage).  FEK forcing after
    the step."""

    def __init__(self, solver, viscos, cfl, closure=None, bf16=True, kf_shell=0, ekinf=None, fused=False):
        self.s, self.viscos, self.cfl = solver, viscos, cfl
        self.closure, self.bf16, self.fused = closure, bf16, fused
        self.kf_shell = kf_shell
        self.kfor = kf_shell + 1 if kf_shell > 0 else 0
        self.ekinf = ekinf if ekinf is not None or not self.kfor else solver.ek_shell(self.kfor)
        self.dts = []

    def step(self, gsh):
        if self.fused:
            return self._step_fused(gsh)
        s, m = self.s, self.closure
        s.set_shifts(gsh)
        for k in (1, 2):
            s.phase1(k)
            if m is not None:
                m.features(s)
                m.predict(self.bf16)
            s.itransform_vel()
            vm = s.realspace_vel(k)
            if m is not None:
                m.add_stress(s)
            s.transform()
            if k == 1:
                self.velmax = vm
                self.dt = self.cfl / vm
                xf, yf, zf = s.step_factors(self.dt)
            s.wavespace_vel(k, self.dt / 2.0)
            if k == 1:
                s.proc4a(xf, yf, zf)
            else:
                s.proc4b()
        if self.kfor:
            s.ek_force(self.kf_shell, self.kfor, self.ekinf)
        self.dts.append(self.dt)

    def _step_fused(self, gsh):
        s, m = self.s, self.closure
        s.set_shifts(gsh)
        for k in (1, 2):
            if m is not None:
                m.features(s, k)
                m.predict(self.bf16)
                m.attach(s, k)
            vm = s.stage_nonlinear(k)
            if k == 1:
                self.velmax = vm
                self.dt = self.cfl / vm
                xf, yf, zf = s.step_factors(self.dt)
            s.stage_update(k, self.dt, xf, yf, zf)
        if self.kfor:
            s.ek_force(self.kf_shell, self.kfor, self.ekinf)
        self.dts.append(self.dt)
