"""ctypes wrapper of the CPU FP64 oracle (oracle/psdns_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The GPU library never uses it.

Arrays follow the reference layouts (0-based, C order):
  cylinder  (nv, num_al, n)  complex128   == uy(ny, num_al, nv)
  physical  (nv, n, n, n)    float64      == ux(nx, zisz, yjsz, nv)  [v][y][z][x]
"""
import ctypes
import os
import pathlib
import subprocess

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
# PSDNS_ORACLE_LIB selects another build of the same sources, e.g. the
# AddressSanitizer/UBSan one of `make -C oracle sanitize`
_LIB_PATH = pathlib.Path(os.environ.get("PSDNS_ORACLE_LIB", _HERE / "_build" / "libpsdns_oracle.so"))
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force=False):
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not _LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(_HERE)] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(_LIB_PATH))
        L.orc_grid_create.restype = ctypes.c_void_p
        L.orc_grid_create.argtypes = [ctypes.c_int]
        L.orc_grid_destroy.argtypes = [ctypes.c_void_p]
        for f in ("orc_num_al", "orc_max_al_x"):
            getattr(L, f).restype = ctypes.c_int
            getattr(L, f).argtypes = [ctypes.c_void_p]
        L.orc_mask_count.restype = ctypes.c_long
        L.orc_mask_count.argtypes = [ctypes.c_void_p]
        L.orc_mask.restype = ctypes.c_int
        L.orc_mask.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_columns.argtypes = [ctypes.c_void_p, _ip, _ip, _ip]
        L.orc_set_shifts.argtypes = [ctypes.c_void_p, _dp]
        L.orc_shift_rule.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.orc_fft_c2c.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_fft_r2c.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_fft_c2r.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_phshift_inplace.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_itransform_vel.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_realspace_vel.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp]
        L.orc_transform.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_wavespace_vel.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_double]
        L.orc_step_factors.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        L.orc_proc4a.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.orc_proc4b.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_kxtran.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.orc_xktran.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.orc_epfor.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int, _dp]
        L.orc_ek_shell.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp]
        L.orc_ek_force.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_spectrum.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_vg_moments.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_field.argtypes = [ctypes.c_void_p, _dp, ctypes.c_ulonglong, ctypes.c_double, ctypes.c_double]
        L.orc_spline3.argtypes = [ctypes.c_int, _dp]
        L.orc_lag_interp.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_long, _dp]
        L.orc_lag_advpos.argtypes = [ctypes.c_int, ctypes.c_long, _dp, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double]
        _lib = L
    return _lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class Grid:
    """Oracle grid for an n^3 box (one task): cylinder, mask, shifts."""

    def __init__(self, n):
        self.n = n
        self.L = lib()
        self.h = self.L.orc_grid_create(n)
        self.num_al = self.L.orc_num_al(self.h)
        self.max_al_x = self.L.orc_max_al_x(self.h)
        cx = np.zeros(self.num_al, np.int32)
        cz = np.zeros(self.num_al, np.int32)
        cut = np.zeros(self.max_al_x, np.int32)
        self.L.orc_columns(self.h, cx.ctypes.data_as(_ip), cz.ctypes.data_as(_ip), cut.ctypes.data_as(_ip))
        self.col_x, self.col_z, self.cut_z = cx, cz, cut

    def __del__(self):
        try:
            self.L.orc_grid_destroy(self.h)
        except Exception:
            pass

    # geometry
    def mask_count(self):
        return self.L.orc_mask_count(self.h)

    def mask(self):
        n = self.n
        kx = self.col_x.astype(np.int64)[:, None]
        kz = np.where(self.col_z > n // 2, self.col_z - n, self.col_z).astype(np.int64)[:, None]
        y = np.arange(n)
        ky = np.where(y > n // 2, y - n, y)[None, :]
        m = (9 * (kx * kx + ky * ky + kz * kz) <= 2 * n * n) & (y[None, :] != n // 2) & (self.col_z[:, None] != n // 2)
        return m

    def set_shifts(self, gsh):
        g = np.ascontiguousarray(np.asarray(gsh, np.float64).reshape(6))
        self.L.orc_set_shifts(self.h, _p(g))

    # fields
    def zeros_cyl(self, nv):
        return np.zeros((nv, self.num_al, self.n), np.complex128)

    def zeros_phys(self, nv):
        return np.zeros((nv, self.n, self.n, self.n), np.float64)

    def field(self, seed=122929298, kpeak=3.0, tke=0.5):
        u = self.zeros_cyl(3)
        self.L.orc_field(self.h, _p(u), seed, kpeak, tke)
        return u

    # hot path
    def phshift_inplace(self, uy, nv, kstep):
        self.L.orc_phshift_inplace(self.h, _p(uy), nv, kstep)

    def itransform_vel(self, uy):
        ux = self.zeros_phys(5)
        self.L.orc_itransform_vel(self.h, _p(np.ascontiguousarray(uy[:3])), _p(ux))
        return ux

    def realspace_vel(self, ux, rkstep):
        vm = np.zeros(1)
        self.L.orc_realspace_vel(self.h, _p(ux), rkstep, _p(vm))
        return float(vm[0])

    def transform(self, ux):
        uy = self.zeros_cyl(5)
        self.L.orc_transform(self.h, _p(ux), _p(uy))
        return uy

    def wavespace_vel(self, uy, kstep, dt2):
        self.L.orc_wavespace_vel(self.h, _p(uy), kstep, dt2)

    def step_factors(self, viscos, dt):
        n = self.n
        xf, yf, zf = np.zeros(n // 2), np.zeros(n), np.zeros(n)
        self.L.orc_step_factors(self.h, viscos, dt, _p(xf), _p(yf), _p(zf))
        return xf, yf, zf

    def proc4a(self, uy, uny, xf, yf, zf):
        self.L.orc_proc4a(self.h, _p(uy), _p(uny), _p(xf), _p(yf), _p(zf))

    def proc4b(self, uy, uny):
        self.L.orc_proc4b(self.h, _p(uy), _p(uny))

    def kxtran(self, uy, nv):
        ux = self.zeros_phys(nv)
        self.L.orc_kxtran(self.h, _p(np.ascontiguousarray(uy[:nv])), _p(ux), nv)
        return ux

    def xktran(self, ux, nv):
        uy = self.zeros_cyl(nv)
        self.L.orc_xktran(self.h, _p(np.ascontiguousarray(ux[:nv])), _p(uy), nv)
        return uy

    def epfor(self, uy, uny, kfor, kforce, forc, mode):
        """EPFOR forcing increments forc (3, k2fo, k2fo, kfor) complex ==
        for(kfor, k2fo, k2fo, 3), in place on uy[:3] (and uny in mode 2):
        mode 1 proc4a.f:588-640 (returns efkz (3, n/2) == efkz(nxh, 3)),
        mode 2 proc4b.f:118-163, mode 3 advanc_2.f/advanc_3.f."""
        assert uy.flags["C_CONTIGUOUS"] and uny.flags["C_CONTIGUOUS"]
        f = np.ascontiguousarray(forc, np.complex128)
        assert f.shape == (3, 2 * kfor - 1, 2 * kfor - 1, kfor)
        efkz = np.zeros((3, self.n // 2))
        self.L.orc_epfor(self.h, _p(uy), _p(uny), kfor, kforce, _p(f), mode, _p(efkz))
        return efkz

    def ek_shell(self, uny, kfor):
        ek = np.zeros(kfor)
        self.L.orc_ek_shell(self.h, _p(np.ascontiguousarray(uny[:3], np.complex128)), kfor, _p(ek))
        return ek

    def ek_force(self, uy, uny, kf_shell, kfor, ekinf):
        e = np.ascontiguousarray(ekinf, np.float64)
        self.L.orc_ek_force(self.h, _p(uy), _p(uny), kf_shell, kfor, _p(e))

    def vg_moments(self, uny):
        """Moments m = 1..8 of du/dx and dw/dz over the grid / n^3
        (PSDNS/vgmomt.f:68-143), from the state on the mask."""
        du, dw = np.zeros(8), np.zeros(8)
        self.L.orc_vg_moments(self.h, _p(np.ascontiguousarray(uny[:3], np.complex128)), _p(du), _p(dw))
        return du, dw

    def spectrum(self, uny, viscos, nshell=None):
        nshell = nshell or self.n
        ek, dk = np.zeros(nshell), np.zeros(nshell)
        t, e = np.zeros(1), np.zeros(1)
        self.L.orc_spectrum(self.h, _p(np.ascontiguousarray(uny[:3], np.complex128)), viscos, nshell, _p(ek), _p(dk), _p(t), _p(e))
        return ek, dk, float(t[0]), float(e[0])


# ---------------------------------------------------------------------------
# passive scalars (numpy over the C transforms)
# ---------------------------------------------------------------------------
def shift_factors(n, gsh):
    """shifts.f:146-178: sx (2, n/2), sy (2, n), sz (2, n) = exp(i 2 pi k
    gsh(dir, stage) / n), k the FFT-order wavenumber."""
    gg = np.asarray(gsh, np.float64).reshape(2, 3)
    kx = np.arange(n // 2, dtype=np.float64)
    k = np.arange(n)
    kw = np.where(k > n // 2, k - n, k).astype(np.float64)
    ph = lambda kk, d: np.exp(1j * (2.0 * np.pi / n * kk[None, :] * gg[:, d:d + 1]))
    return ph(kx, 0), ph(kw, 1), ph(kw, 2)


def scalar_shift(g, gsh, kstep):
    """sx(kx) sz(kz) sy(ky) of stage kstep on the cylinder (num_al, n)"""
    sx, sy, sz = shift_factors(g.n, gsh)
    s = kstep - 1
    return (sx[s][g.col_x] * sz[s][g.col_z])[:, None] * sy[s][None, :]


def scalar_terms(g, ux, phi, kstep, grad, gsh):
    """Advection terms of the passive scalars phi (nc, num_al, n), stage
    kstep, with ux[0:3] the phase-shifted physical velocity:
      phase1.f:44-64          phi_s = sx sy sz phi on the mask (0 elsewhere)
      phase1.f:58, itransform.f:190-203
                              d phi/dy, d phi/dz, d phi/dx = i k phi_s
      itransform.f:522-531    term = u (G1 + phi_x) + v (G2 + phi_y) + w (G3 + phi_z)
      transform.f             forward transform of term; here fully
                              normalised (xktran: 1/(nx ny nz)), which is the
                              reference's x r2c / nx with wavespace's
                              1/(ny nz) (wavespace_module.f:347-352)
    Returns xktran(term) (nc, num_al, n)."""
    n = g.n
    m = g.mask()
    kx = g.col_x.astype(np.float64)[:, None]
    kz = np.where(g.col_z > n // 2, g.col_z - n, g.col_z).astype(np.float64)[:, None]
    y = np.arange(n)
    ky = np.where(y > n // 2, y - n, y).astype(np.float64)[None, :]
    sh = scalar_shift(g, gsh, kstep)
    out = g.zeros_cyl(len(phi))
    for i in range(len(phi)):
        ps = np.where(m, sh * phi[i], 0)
        d = g.zeros_cyl(3)
        d[0], d[1], d[2] = 1j * kx * ps, 1j * ky * ps, 1j * kz * ps
        f = g.kxtran(d, 3)
        term = ux[0] * (grad[i][0] + f[0]) + ux[1] * (grad[i][1] + f[1]) + ux[2] * (grad[i][2] + f[2])
        out[i] = g.xktran(np.ascontiguousarray(term[None]), 1)[0]
    return out


def scalar_update(g, phi_u, phi_n, terms, kstep, dt, gsh, factors=None, pr=None):
    """Scalar part of the update, in place on phi_u / phi_n (nc, num_al, n):
      wavespace_module.f:347-352   N = -(dt/2) conj(sx sy sz) terms on the mask
      proc4a.f:506-520 (stage 1)   F = xdif ydif zdif, xdif = xfac^(1/pr)
                                   (step.f:387-407);
                                   phi_n <- F (phi_n + N), phi_u <- F (phi_n + 2 N)
      proc4b.f:69-74, 183-190, 201-205 (stage 2)
                                   phi_u <- phi_n + N, phi_n <- phi_u, zero mean"""
    m = g.mask()
    nsh = -(dt / 2.0) * np.conj(scalar_shift(g, gsh, kstep))
    for i in range(len(phi_u)):
        nl = nsh * terms[i]
        if kstep == 1:
            xf, yf, zf = factors
            r = 1.0 / pr[i]
            f = (xf[g.col_x] ** r * zf[g.col_z] ** r)[:, None] * (yf ** r)[None, :]
            un = phi_n[i].copy()
            phi_n[i] = np.where(m, f * (un + nl), phi_n[i])
            phi_u[i] = np.where(m, f * (un + 2.0 * nl), phi_u[i])
        else:
            phi_u[i] = np.where(m, phi_n[i] + nl, phi_u[i])
            phi_n[i] = np.where(m, phi_u[i], phi_n[i])
            phi_u[i, 0, 0] = 0
            phi_n[i, 0, 0] = 0


def scalar_advanc(g, k, phi_u, phi_n, phi_1, terms, ks, dt, gsh, factors=None, pr=None):
    """Scalar loops of the RK4 substep k (i = 4 .. 3 + nc of
    PSDNS/homogeneous/advanc_1.f:101-111, advanc_2.f, advanc_3.f,
    advanc_4.f:46-52,149-176), in place on phi_u, phi_n and the accumulator
    phi_1, with N = -(dt/2) conj(sx sy sz) terms (wavespace_module.f:347-352,
    shifts of stage ks) on the mask and, for k = 1, the factor
    xdif ydif zdif with xdif = xfac^(1/pr) (step.f:387-407).  No
    projection; zero mean after k = 4.  Pinned like advanc (the same weights
    and placement, checked through the velocity)."""
    m = g.mask()
    nsh = -(dt / 2.0) * np.conj(scalar_shift(g, gsh, ks))
    for i in range(len(phi_u)):
        nl = nsh * terms[i]
        if k == 1:
            xf, yf, zf = factors
            r = 1.0 / pr[i]
            f = (xf[g.col_x] ** r * zf[g.col_z] ** r)[:, None] * (yf ** r)[None, :]
            a = f * (phi_n[i] + nl / 3.0)
            u = f * (phi_n[i] + nl)
        elif k in (2, 3):
            a = phi_1[i] + nl * 2.0 / 3.0
            u = phi_n[i] + nl * (1.0 if k == 2 else 2.0)
        else:
            a = phi_1[i] + nl * 1.0 / 3.0
            a[0, 0] = 0
            u = a
        phi_1[i] = np.where(m, a, phi_1[i])
        phi_u[i] = np.where(m, u, phi_u[i])
        if k == 4:
            phi_n[i] = np.where(m, a, phi_n[i])


def advanc(g, k, uy, uny, u1, factors=None):
    """RK4 substep k = 1..4 (rkmethod = 4) on the cylinder, in place on
    uy[:3] (the nonlinear term N on entry, dt/2 attached), uny and u1, on
    the mask only.  Restates PSDNS/homogeneous/advanc_1.f:88-153 (the
    integrating factor xfac(x) yfac(y) zfac(z) and the modpress projection,
    k = 0 skipped), advanc_2.f and advanc_3.f (u1 += N 2/3; u = un + N, resp.
    un + 2 N; projection) and advanc_4.f:46-83,149-176 (u1 += N/3; projection
    of u1; un = u = u1 on the mask; zero mean).  EPFOR forcing is not part of
    it.  factors = (xfac, yfac, zfac) for k = 1.

    Pinned by the exact Stokes decay of a Beltrami field (the factor on
    un and N/3 in u1), fourth order in dt at nu = 0 (the substep weights)
    and first order at nu > 0 (no factor after substep 1).  Parity unpinned:
    the factor on substep 1's intermediate u (u = F (un + N)) is fixed by
    the reference text alone; none of these checks would fail without it."""
    n = g.n
    m = g.mask()
    kx = g.col_x.astype(np.float64)[:, None]
    kz = np.where(g.col_z > n // 2, g.col_z - n, g.col_z).astype(np.float64)[:, None]
    y = np.arange(n)
    ky = np.where(y > n // 2, y - n, y).astype(np.float64)[None, :]
    k2 = kx * kx + ky * ky + kz * kz
    k0 = k2 == 0
    kk = (np.broadcast_to(kx, k2.shape), np.broadcast_to(ky, k2.shape), np.broadcast_to(kz, k2.shape))

    def project(v):
        psi = -(kk[0] * v[0] + kk[1] * v[1] + kk[2] * v[2]) / np.where(k0, 1.0, k2)
        psi[k0] = 0
        return np.array([v[i] + kk[i] * psi for i in range(3)])

    nl = uy[:3]
    if k == 1:
        xf, yf, zf = factors
        f = xf[g.col_x][:, None] * yf[None, :] * zf[g.col_z][:, None]
        a = f * (uny + nl / 3.0)
        u = project(f * (uny + nl))
    elif k in (2, 3):
        a = u1 + nl * 2.0 / 3.0
        u = project(uny + nl * (1.0 if k == 2 else 2.0))
    else:
        a = project(u1 + nl * 1.0 / 3.0)
        a[:, k0] = 0
        u = a
    u1[:, m] = a[:, m]
    uy[:3][:, m] = u[:, m]
    if k == 4:
        uny[:, m] = a[:, m]


def epfor_increments(kfor, kforce, rng, amp=1.0):
    """Forcing increments for(kfor, k2fo, k2fo, 3) as force.f:70-160 forms
    them from two complex random amplitudes per node (alpha, beta: the
    reference's Uhlenbeck-Ornstein velinc, here Gaussian draws scaled by
    amp): zero outside |k| <= kforce and at k = 0; (alpha, beta, 0) on the
    kz axis (k12 = 0); otherwise alpha normal to k and z, beta normal to
    alpha and k (solenoidal); conjugate symmetry on kx = 0 (force.f:140-156).
    Returned as a (3, k2fo, k2fo, kfor) complex array (C order of the
    Fortran array)."""
    k2fo = 2 * kfor - 1
    f = np.zeros((3, k2fo, k2fo, kfor), np.complex128)
    vel = amp * (rng.standard_normal((2, k2fo, k2fo, kfor)) + 1j * rng.standard_normal((2, k2fo, k2fo, kfor)))

    def wave(i):  # 1-based index -> wavenumber (force.f ifwave with ky/kz of waveno.f)
        return i - 1 if i <= kfor else i - 2 * kfor

    for z in range(1, k2fo + 1):
        k3 = float(wave(z))
        for y in range(1, k2fo + 1):
            k2 = float(wave(y))
            for x in range(1, kfor + 1):
                al, be = vel[0, z - 1, y - 1, x - 1], vel[1, z - 1, y - 1, x - 1]
                k1 = float(x - 1)
                k = np.sqrt(k1 * k1 + k2 * k2 + k3 * k3)
                k12 = np.sqrt(k1 * k1 + k2 * k2)
                if k > kforce or k == 0.0:
                    continue
                if k12 == 0.0:
                    f[0, z - 1, y - 1, x - 1] = al
                    f[1, z - 1, y - 1, x - 1] = be
                else:
                    f[0, z - 1, y - 1, x - 1] = (al * k * k2 + be * k1 * k3) / (k * k12)
                    f[1, z - 1, y - 1, x - 1] = (be * k2 * k3 - al * k * k1) / (k * k12)
                    f[2, z - 1, y - 1, x - 1] = -be * k12 / k
    for i in range(3):
        for y in range(kfor + 1, k2fo + 1):
            f[i, 0, y - 1, 0] = np.conj(f[i, 0, k2fo + 1 - y, 0])
        for z in range(kfor + 1, k2fo + 1):
            f[i, z - 1, 0, 0] = np.conj(f[i, k2fo + 1 - z, 0, 0])
            for y in range(2, k2fo + 1):
                f[i, z - 1, y - 1, 0] = np.conj(f[i, k2fo + 1 - z, k2fo + 1 - y, 0])
    return f


# ---------------------------------------------------------------------------
# Lagrangian particles (hooks/LAG_LOCAL): pp is the reference's pp(np, 9),
# column-major, as a (9, np) float64 array: rows 0-2 positions in grid units
# (gx(i) = 1 + (i - 1) + gsh(1,1)), 3-5 velocities, 6-8 rk_advpos's saved
# half-step positions
# ---------------------------------------------------------------------------
def spline3(f):
    """spxyz_m.f: tricubic-spline coefficients of periodic fields f[..., y, z, x]"""
    out = np.ascontiguousarray(f, np.float64).copy()
    n = out.shape[-1]
    for v in out.reshape(-1, n, n, n):
        lib().orc_spline3(n, _p(v))
    return out


def velocity_spline(g, uny, gsh):
    """rk_part.f:28-41: spline coefficients (3, n, n, n) of the velocity of
    uny on the stage-1 phase-shifted grid (the physical u of itransform_vel
    at kstep 1)"""
    uy = g.zeros_cyl(5)
    uy[:3] = uny[:3]
    g.set_shifts(gsh)
    g.phshift_inplace(uy, 3, 1)
    return spline3(g.kxtran(uy, 3))


def lag_interp(c, gsh1, pp):
    """spcal_local.f: pp[3:6] = spline velocity at pp[0:3] (in place)"""
    assert pp.flags["C_CONTIGUOUS"] and pp.shape[0] == 9
    c = np.ascontiguousarray(c, np.float64)
    g3 = np.ascontiguousarray(np.asarray(gsh1, np.float64)[:3])
    lib().orc_lag_interp(c.shape[-1], _p(c), c.shape[0], _p(g3), pp.shape[1], _p(pp))


def lag_advpos(n, pp, pstep, dt, cof1, cof2):
    """rk_advpos.f (in place)"""
    assert pp.flags["C_CONTIGUOUS"] and pp.shape[0] == 9
    lib().orc_lag_advpos(n, pp.shape[1], _p(pp), pstep, dt, cof1, cof2)


def particles_uniform(n, np_, seed):
    """np_ particles uniform in the box: positions 1 + U n in grid units"""
    pp = np.zeros((9, np_))
    pp[:3] = 1.0 + np.random.default_rng(seed).random((3, np_)) * n
    return pp


def fft_c2c(x, sign):
    a = np.ascontiguousarray(x, np.complex128).copy()
    lib().orc_fft_c2c(_p(a), a.shape[-1], a.size // a.shape[-1], sign)
    return a


def fft_r2c(x):
    x = np.ascontiguousarray(x, np.float64)
    n = x.shape[-1]
    out = np.zeros(x.shape[:-1] + (n // 2 + 1,), np.complex128)
    lib().orc_fft_r2c(_p(x), _p(out), n, x.size // n)
    return out


def fft_c2r(X, n):
    X = np.ascontiguousarray(X, np.complex128)
    out = np.zeros(X.shape[:-1] + (n,), np.float64)
    lib().orc_fft_c2r(_p(X), _p(out), n, X.size // X.shape[-1])
    return out


class Run:
    """RK2 time loop on the oracle, mirroring main.f/timestep_module.f:
    shifts -> [phase1, itransform_vel, realspace_vel, transform,
    (stage 1: step -> dt, factors), wavespace_vel, proc4a | proc4b] x 2,
    then FEK forcing (ek_force) after the corrector."""

    def __init__(self, grid, uny, viscos, cfl, kshift=4, kf_shell=0, shift_seed=7, ekinf=None, dt_fixed=None,
                 phi=None, pr=None, grad=None, epfor=None, particles=None):
        """epfor = (kfor, kforce, forcing) with forcing(istep) the increments
        for(kfor, k2fo, k2fo, 3) of step istep (0-based), (3, k2fo, k2fo,
        kfor) complex: EPFOR forcing after proc4a/proc4b (advanc_1..3)."""
        self.g = grid
        self.epfor = epfor
        # Lagrangian particles pp (9, np), advanced at the start of each step
        # (lagpart from realspace at kstep 1, before step: the previous dt)
        self.pp = None if particles is None else np.ascontiguousarray(particles, np.float64).copy()
        self.efkz = None
        self.uny = np.ascontiguousarray(uny, np.complex128).copy()
        self.uy = grid.zeros_cyl(5)
        self.uy[:3] = self.uny
        # passive scalars: state phi_n and the stage value phi_u
        self.phi_n = None if phi is None else np.ascontiguousarray(phi, np.complex128).copy()
        self.phi_u = None if phi is None else self.phi_n.copy()
        self.pr, self.grad = pr, grad
        self.viscos, self.cfl, self.kshift = viscos, cfl, kshift
        self.dt_fixed = dt_fixed
        self.kf_shell = kf_shell
        self.kfor = kf_shell + 1 if kf_shell > 0 else 0
        self.ekinf = ekinf if ekinf is not None else (grid.ek_shell(self.uny, self.kfor) if self.kfor else None)
        self.gsh = np.zeros(6)
        self.rng = np.random.default_rng(shift_seed)
        self.istep = 0
        self.time = 0.0
        self.dts = []
        if kshift in (1, 3):
            r = np.ascontiguousarray(self.rng.random(3))
            lib().orc_shift_rule(kshift, _p(self.gsh), _p(r), 1)

    def shift_sequence(self, nsteps):
        """the gsh(3,2) values the next nsteps steps will use (consumes them)"""
        return np.array([self.next_shifts() for _ in range(nsteps)])

    def next_shifts(self):
        r = np.ascontiguousarray(self.rng.random(3))
        lib().orc_shift_rule(self.kshift, _p(self.gsh), _p(r), 0)
        return self.gsh.copy()

    def particles_step(self, gsh_now):
        """partic(4) on the first step (interpolation only), else partsp's
        RK2: rk_advpos(1, 1/2, 1/2), the spline of the state on this step's
        stage-1 grid and the interpolation at the predicted positions,
        rk_advpos(2, 0, 1/2), all with the previous step's dt
        (partsp.f:217-299)"""
        g, pp = self.g, self.pp
        c = velocity_spline(g, self.uny, gsh_now)
        g.set_shifts(gsh_now)
        if not self.dts:
            lag_interp(c, gsh_now[:3], pp)
            return
        dt = self.dts[-1]
        lag_advpos(g.n, pp, 1, dt, 0.5, 0.5)
        lag_interp(c, gsh_now[:3], pp)
        lag_advpos(g.n, pp, 2, dt, 0.0, 0.5)

    def step(self, gsh=None):
        g = self.g
        g.set_shifts(self.next_shifts() if gsh is None else gsh)
        gsh_now = self.gsh if gsh is None else np.asarray(gsh)
        if self.pp is not None:
            self.particles_step(np.asarray(gsh_now, np.float64).reshape(6))
        for k in (1, 2):
            g.phshift_inplace(self.uy, 3, k)
            ux = g.itransform_vel(self.uy)
            if self.phi_n is not None:
                terms = scalar_terms(g, ux, self.phi_u, k, self.grad, gsh_now)
            vm = g.realspace_vel(ux, k)
            nl = g.transform(ux)
            if k == 1:
                self.velmax = vm
                # step.f:77-78 (cfl > 0), else a constant dt
                self.dt = self.cfl / vm if self.dt_fixed is None else self.dt_fixed
                self.factors = g.step_factors(self.viscos, self.dt)
            g.wavespace_vel(nl, k, self.dt / 2.0)
            self.uy = nl
            if k == 1:
                g.proc4a(self.uy, self.uny, *self.factors)
            else:
                g.proc4b(self.uy, self.uny)
            if self.epfor is not None:
                kf, kfo, fn = self.epfor
                e = g.epfor(self.uy, self.uny, kf, kfo, fn(self.istep), k)
                if k == 1:
                    self.efkz = e
            if self.phi_n is not None:
                scalar_update(g, self.phi_u, self.phi_n, terms, k, self.dt, gsh_now, self.factors, self.pr)
        if self.kfor:
            g.ek_force(self.uy, self.uny, self.kf_shell, self.kfor, self.ekinf)
        self.istep += 1
        self.time += self.dt
        self.dts.append(self.dt)
        return self.uny

    def step_rk4(self, gsh=None):
        """RK4 step (rkmethod = 4, PSDNS/timestep_module.f:293-310): four
        substeps through the same nonlinear term, stages 3 and 4 with the
        shifts of stages 1 and 2 (PSDNS/shifts.f:133-137), dt and the
        integrating factors from substep 1, advanc_1..4, then FEK forcing.
        Passive scalars (phi) go through scalar_terms and scalar_advanc."""
        g = self.g
        g.set_shifts(self.next_shifts() if gsh is None else gsh)
        gsh_now = self.gsh if gsh is None else np.asarray(gsh)
        if not hasattr(self, "u1"):
            self.u1 = g.zeros_cyl(3)  # comp_set.f:72-74
            if self.phi_n is not None:
                self.phi_1 = np.zeros_like(self.phi_n)
        for k in (1, 2, 3, 4):
            ks = 1 if k in (1, 3) else 2
            g.phshift_inplace(self.uy, 3, ks)
            ux = g.itransform_vel(self.uy)
            if self.phi_n is not None:
                terms = scalar_terms(g, ux, self.phi_u, ks, self.grad, gsh_now)
            vm = g.realspace_vel(ux, 1 if k == 1 else 2)
            nl = g.transform(ux)
            if k == 1:
                self.velmax = vm
                self.dt = self.cfl / vm if self.dt_fixed is None else self.dt_fixed
                self.factors = g.step_factors(self.viscos, self.dt)
            g.wavespace_vel(nl, ks, self.dt / 2.0)
            self.uy = nl
            advanc(g, k, self.uy, self.uny, self.u1, self.factors)
            if self.epfor is not None and k < 4:
                # advanc_1 (tally) .. advanc_3 add to u; advanc_4's add is
                # overwritten by u1 on the mask (advanc_4.f:96-170)
                kf, kfo, fn = self.epfor
                e = g.epfor(self.uy, self.uny, kf, kfo, fn(self.istep), 1 if k == 1 else 3)
                if k == 1:
                    self.efkz = e
            if self.phi_n is not None:
                scalar_advanc(g, k, self.phi_u, self.phi_n, self.phi_1, terms, ks, self.dt, gsh_now, self.factors,
                              self.pr)
        if self.kfor:
            g.ek_force(self.uy, self.uny, self.kf_shell, self.kfor, self.ekinf)
        self.istep += 1
        self.time += self.dt
        self.dts.append(self.dt)
        return self.uny
