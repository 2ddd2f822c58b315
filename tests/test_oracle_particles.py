"""Oracle pins (CPU) for the Lagrangian-particle path (hooks/LAG_LOCAL):
- node exactness, the reference's own check (pitest.f): the spline
  interpolant at a grid node of the shifted grid returns the grid value;
- the solve is the periodic (1,4,1) system with the splz factor 216: the
  coefficients equal the exact spectral inverse 216 / prod(4 + 2 cos theta)
  computed with numpy's FFT, and a constant field has constant coefficients
  equal to it (partition of unity);
- interpolation of a smooth field converges at fourth order in h;
- rk_advpos in a frozen uniform-along-the-path flow is exact (Heun)."""
import numpy as np
import pytest

from conftest import rel


def grid_xyz(n):
    x = np.arange(n) * 2 * np.pi / n
    return np.meshgrid(x, x, x, indexing="ij")  # Y, Z, X for [y][z][x]


def test_spline_is_the_exact_periodic_inverse(O):
    import oracle as orc
    n = 16
    f = np.random.default_rng(1).standard_normal((n, n, n))
    c = orc.spline3(f)
    th = 2 * np.pi * np.fft.fftfreq(n) * n / n
    lam = 4 + 2 * np.cos(2 * np.pi * np.arange(n) / n)
    F = np.fft.fftn(f)
    ref = np.real(np.fft.ifftn(F * 216 / (lam[:, None, None] * lam[None, :, None] * lam[None, None, :])))
    assert rel(c, ref) < 1e-14
    one = orc.spline3(np.full((n, n, n), 2.5))
    assert np.allclose(one, 2.5, rtol=1e-14, atol=0)


@pytest.mark.parametrize("gsh", [(0.0, 0.0, 0.0), (0.3, 0.71, 0.45)])
def test_node_exactness(O, gsh):
    import oracle as orc
    n = 16
    f = np.random.default_rng(2).standard_normal((1, n, n, n))
    c = orc.spline3(f)
    rng = np.random.default_rng(3)
    idx = rng.integers(0, n, (3, 200))  # (x, y, z) node indices
    pp = np.zeros((9, 200))
    for j in range(3):
        pp[j] = 1.0 + idx[j] + gsh[j]  # gx(i) = 1 + (i - 1) + gsh
    # nodes wrapped around the box by whole periods land on the same nodes
    pp[0, :50] += n
    pp[1, 50:100] -= 2 * n
    orc.lag_interp(c, gsh, pp)
    want = f[0, idx[1], idx[2], idx[0]]
    assert np.abs(pp[3] - want).max() < 1e-13 * np.abs(f).max()


def test_fourth_order_interpolation(O):
    import oracle as orc
    errs = []
    rng = np.random.default_rng(4)
    pts = rng.random((3, 300)) * 2 * np.pi
    fn = lambda X, Y, Z: np.sin(X + 2 * Y) * np.cos(Z) + 0.5 * np.cos(2 * X - Y + Z)
    for n in (16, 32, 64):
        Y, Z, X = grid_xyz(n)
        c = orc.spline3(fn(X, Y, Z)[None])
        pp = np.zeros((9, pts.shape[1]))
        pp[:3] = 1.0 + pts * n / (2 * np.pi)
        orc.lag_interp(c, (0, 0, 0), pp)
        errs.append(np.abs(pp[3] - fn(pts[0], pts[1], pts[2])).max())
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 12 < r1 < 20 and 12 < r2 < 20, errs


def test_rk_advpos_heun_frozen_shear(O):
    """u = (A sin y, 0, 0) frozen: the velocity is constant along a particle
    path, so x(dt) = x + A sin(y) dt exactly; y, z unchanged"""
    import oracle as orc
    n, A, dt = 32, 0.8, 0.05
    Y, Z, X = grid_xyz(n)
    c = orc.spline3(np.stack([A * np.sin(Y), 0 * Y, 0 * Y]))
    pp = orc.particles_uniform(n, 100, 5)
    x0 = pp[:3].copy()
    orc.lag_interp(c, (0, 0, 0), pp)
    u0 = pp[3].copy()
    orc.lag_advpos(n, pp, 1, dt, 0.5, 0.5)
    orc.lag_interp(c, (0, 0, 0), pp)
    orc.lag_advpos(n, pp, 2, dt, 0.0, 0.5)
    dx = 2 * np.pi / n
    assert np.allclose(pp[0], x0[0] + u0 * dt / dx, rtol=0, atol=1e-12)
    assert np.allclose(pp[1:3], x0[1:3], rtol=0, atol=1e-12)
    # the interpolated u0 is the analytic field to spline accuracy
    yphys = (x0[1] - 1) * dx
    assert np.abs(u0 - A * np.sin(yphys)).max() < 1e-4


def test_particles_in_the_run_follow_the_field(O):
    """a shear flow u = (A sin y, 0, 0) decaying as exp(-nu t): x advances by
    the time integral of the interpolated velocity, to second order"""
    import oracle as orc
    n, A, nu = 16, 1.0, 0.05
    g = O.Grid(n)
    u0 = g.zeros_cyl(3)
    a = int(np.nonzero((g.col_x == 0) & (g.col_z == 0))[0][0])
    u0[0, a, 1], u0[0, a, n - 1] = A / 2j, -A / 2j
    pp = orc.particles_uniform(n, 50, 6)
    x0 = pp[:3].copy()
    r = O.Run(g, u0, nu, cfl=-1.0, kshift=4, dt_fixed=0.05, particles=pp)
    for _ in range(5):
        r.step()
    # after 5 steps the particles have moved through t = 4 dt (one step lag)
    T = 4 * 0.05
    dx = 2 * np.pi / n
    y = (x0[1] - 1) * dx
    want = x0[0] + A * np.sin(y) * (1 - np.exp(-nu * T)) / nu / dx
    assert np.abs(r.pp[0] - want).max() < 2e-3
    assert np.allclose(r.pp[1:3], x0[1:3], atol=1e-12)
