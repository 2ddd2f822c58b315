"""Oracle pins (CPU): transforms, nonlinear term, update, RK2.

Each check is independent of the oracle's own formulas:
- kxtran/xktran vs numpy irfftn/rfftn on the same cylinder data;
- phase shift vs the shift theorem (a Fourier mode evaluated on the shifted grid);
- realspace products pointwise;
- the full nonlinear chain (phase1 -> itransform_vel -> realspace_vel ->
  transform -> wavespace_vel) vs a numpy pseudo-spectral evaluation of
  -(dt/2) F[d_j(u_i u_j) - d_i(v v)] for band-limited fields (no aliasing);
- energy conservation of the projected band-limited nonlinear term;
- proc4a/proc4b projection: k.u = 0 and idempotence;
- Stokes decay of an ABC (Beltrami) field: u(t) = exp(-nu k^2 t) u(0);
- second-order convergence in dt; the reference's own invariants for the
  sinusoidal IC (README.spp:101-105: TKE does not increase, rms u = rms v).
"""
import numpy as np
import pytest

from conftest import kgrid, random_cylinder, rel


def to_full(grid, uy):
    """cylinder (nv, num_al, n) -> numpy rfftn layout (nv, n_y, n_z, n/2+1)"""
    n = grid.n
    H = np.zeros((uy.shape[0], n, n, n // 2 + 1), complex)
    for a in range(grid.num_al):
        H[:, :, grid.col_z[a], grid.col_x[a]] = uy[:, a, :]
    return H


def from_full(grid, H):
    out = np.zeros((H.shape[0], grid.num_al, grid.n), complex)
    for a in range(grid.num_al):
        out[:, a, :] = H[:, :, grid.col_z[a], grid.col_x[a]]
    return out


@pytest.mark.parametrize("n", [16, 32])
def test_kxtran_vs_numpy(O, n):
    g = O.Grid(n)
    u = random_cylinder(g, 3, 1)
    ux = g.kxtran(u, 3)
    H = to_full(g, u)
    H[:, n // 2] = 0  # kxtran.f:36
    ref = np.fft.irfftn(H, s=(n, n, n), axes=(1, 2, 3)) * n ** 3
    assert rel(ux, ref) < 1e-13


@pytest.mark.parametrize("n", [16, 32])
def test_xktran_vs_numpy(O, n):
    g = O.Grid(n)
    rng = np.random.default_rng(5)
    ux = rng.standard_normal((2, n, n, n))
    uy = g.xktran(ux, 2)
    ref = from_full(g, np.fft.rfftn(ux, axes=(1, 2, 3)) / n ** 3)
    assert rel(uy, ref) < 1e-13


def test_round_trip_is_identity_on_the_mask(O):
    g = O.Grid(32)
    u = random_cylinder(g, 3, 2)
    back = g.xktran(g.kxtran(u, 3), 3)
    assert rel(back, u, g.mask()) < 1e-14


def test_single_mode_and_shift_theorem(O):
    n = 32
    g = O.Grid(n)
    kx, ky, kz = 3, -5, 2
    a = next(i for i in range(g.num_al) if g.col_x[i] == kx and g.col_z[i] == kz % n)
    u = g.zeros_cyl(5)
    u[0, a, ky % n] = 0.5 - 0.25j
    gsh = np.array([0.3, 0.7, 0.1, 0.0, 0.0, 0.0])
    g.set_shifts(gsh)
    g.phshift_inplace(u, 3, 1)
    ux = g.itransform_vel(u)[0]
    x = np.arange(n) * 2 * np.pi / n
    dx = 2 * np.pi / n
    Y, Z, X = np.meshgrid(x + gsh[1] * dx, x + gsh[2] * dx, x + gsh[0] * dx, indexing="ij")
    ref = 2 * np.real((0.5 - 0.25j) * np.exp(1j * (kx * X + ky * Y + kz * Z)))
    assert np.abs(ux - ref).max() < 1e-13


def test_realspace_products_pointwise(O):
    n = 16
    g = O.Grid(n)
    rng = np.random.default_rng(9)
    ux = np.zeros((5, n, n, n))
    ux[:3] = rng.standard_normal((3, n, n, n))
    u, v, w = ux[0].copy(), ux[1].copy(), ux[2].copy()
    vm = g.realspace_vel(ux, 1)
    assert np.allclose(ux[0], u * u - v * v, rtol=0, atol=1e-15 * 10)
    assert np.allclose(ux[1], u * v, rtol=0, atol=1e-15)
    assert np.allclose(ux[2], w * w - v * v, rtol=0, atol=1e-14)
    assert np.allclose(ux[3], v * w, rtol=0, atol=1e-15)
    assert np.allclose(ux[4], u * w, rtol=0, atol=1e-15)
    assert vm == pytest.approx(n * np.max(np.abs(u) + np.abs(v) + np.abs(w)), rel=1e-15)


def band_limited(g, seed, kcut):
    u = random_cylinder(g, 3, seed)
    kx, ky, kz = kgrid(g.n, g.col_x, g.col_z)
    keep = (kx * kx + ky * ky + kz * kz) < kcut * kcut
    return np.ascontiguousarray(u * keep[None])


def numpy_nonlinear(g, u, dt2):
    """-(dt/2) F[d_j(u_i u_j) - d_i(v v)] on the cylinder, by numpy"""
    n = g.n
    H = to_full(g, u)
    phys = np.fft.irfftn(H, s=(n, n, n), axes=(1, 2, 3)) * n ** 3
    U, V, Wv = phys
    k = np.fft.fftfreq(n, 1.0 / n)
    KY = k[:, None, None]
    KZ = k[None, :, None]
    KX = np.arange(n // 2 + 1)[None, None, :]
    F = lambda f: np.fft.rfftn(f) / n ** 3  # noqa: E731
    dx = lambda f: 1j * KX * F(f)  # noqa: E731
    dy = lambda f: 1j * KY * F(f)  # noqa: E731
    dz = lambda f: 1j * KZ * F(f)  # noqa: E731
    N1 = dx(U * U) + dy(U * V) + dz(U * Wv) - dx(V * V)
    N2 = dx(V * U) + dy(V * V) + dz(V * Wv) - dy(V * V)
    N3 = dx(Wv * U) + dy(Wv * V) + dz(Wv * Wv) - dz(V * V)
    return -dt2 * from_full(g, np.stack([N1, N2, N3]))


def oracle_nonlinear(g, u, kstep, dt2):
    uy = g.zeros_cyl(5)
    uy[:3] = u
    g.phshift_inplace(uy, 3, kstep)
    ux = g.itransform_vel(uy)
    g.realspace_vel(ux, kstep)
    nl = g.transform(ux)
    g.wavespace_vel(nl, kstep, dt2)
    return nl


@pytest.mark.parametrize("gsh", [np.zeros(6), np.array([0.2, 0.9, 0.4, 0.6, 0.1, 0.35])])
def test_nonlinear_term_vs_numpy(O, gsh):
    n = 32
    g = O.Grid(n)
    g.set_shifts(gsh)
    u = band_limited(g, 4, n / 4.0 - 0.5)  # products alias-free on the grid
    dt2 = 0.0125
    for kstep in (1, 2):
        nl = oracle_nonlinear(g, u, kstep, dt2)
        ref = numpy_nonlinear(g, u, dt2)
        assert rel(nl[:3], ref, g.mask()[None]) < 1e-12


def test_nonlinear_term_conserves_energy(O):
    n = 32
    g = O.Grid(n)
    g.set_shifts(np.array([0.3, 0.3, 0.3, 0.8, 0.8, 0.8]))
    ones = [np.ones(n // 2), np.ones(n), np.ones(n)]
    # solenoidal band-limited u: proc4a with F = 1, un = 0 gives u <- P(2 u)
    w = g.zeros_cyl(5)
    w[:3] = band_limited(g, 6, n / 4.0 - 0.5)
    g.proc4a(w, g.zeros_cyl(3), *ones)
    u = np.ascontiguousarray(w[:3] / 2.0)
    nl = oracle_nonlinear(g, u, 1, 0.01)
    un = g.zeros_cyl(3)
    g.proc4a(nl, un, *ones)
    pn = nl[:3] / 2.0
    tf = np.where(g.col_x == 0, 1.0, 2.0)[None, :, None]
    power = np.sum(tf * np.real(np.conj(u) * pn))
    scale = np.sum(tf * np.abs(u) * np.abs(pn))
    assert abs(power) < 1e-12 * scale


def test_projection(O):
    n = 32
    g = O.Grid(n)
    uy = g.zeros_cyl(5)
    uy[:3] = random_cylinder(g, 3, 8)
    un = g.zeros_cyl(3)
    ones = [np.ones(n // 2), np.ones(n), np.ones(n)]
    g.proc4a(uy, un, *ones)  # uy <- P(2 uy), un <- uy
    kx, ky, kz = kgrid(n, g.col_x, g.col_z)
    div = kx * uy[0] + ky * uy[1] + kz * uy[2]
    assert np.abs(div).max() < 1e-13 * np.abs(uy).max()
    before = uy[:3].copy()
    before[:, 0, 0] = 0  # proc4b also removes the mean
    un2 = g.zeros_cyl(3)
    g.proc4b(uy, un2)  # uy <- P(0 + uy): idempotent
    assert rel(uy[:3], before) < 1e-15
    assert np.abs(uy[:, 0, 0]).max() == 0.0  # zero mean (proc4b.f:200-205)


def abc_field(g, A=1.0, B=0.7, C=0.4):
    """u = (A sin z + C cos y, B sin x + A cos z, C sin y + B cos x): curl u = u"""
    n = g.n
    x = np.arange(n) * 2 * np.pi / n
    Y, Z, X = np.meshgrid(x, x, x, indexing="ij")
    ux = np.stack([A * np.sin(Z) + C * np.cos(Y), B * np.sin(X) + A * np.cos(Z), C * np.sin(Y) + B * np.cos(X)])
    return g.xktran(ux, 3)


def test_stokes_decay_of_beltrami_field(O):
    g = O.Grid(32)
    u0 = abc_field(g)
    nu = 0.05
    r = O.Run(g, u0, nu, cfl=-1.0, kshift=1, dt_fixed=0.02)
    for _ in range(5):
        r.step()
    assert rel(r.uny, u0 * np.exp(-nu * r.time), g.mask()[None]) < 1e-12


def tg_field(g):
    n = g.n
    x = np.arange(n) * 2 * np.pi / n
    Y, Z, X = np.meshgrid(x, x, x, indexing="ij")
    ux = np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X])
    return g.xktran(ux, 3)


def test_second_order_in_dt(O):
    g = O.Grid(32)
    u0 = tg_field(g)
    nu, T = 0.05, 0.4

    def run(dt):
        r = O.Run(g, u0, nu, cfl=-1.0, kshift=4, dt_fixed=dt)
        for _ in range(int(round(T / dt))):
            r.step()
        return r.uny

    ref = run(0.0125)
    e1 = np.linalg.norm((run(0.1) - ref).ravel())
    e2 = np.linalg.norm((run(0.05) - ref).ravel())
    assert 3.3 < e1 / e2 < 4.7


def test_rk4_stokes_decay_of_beltrami_field(O):
    """The reference's RK4 (advanc_1..4) applies the integrating factor in
    substep 1 only; for a Beltrami field the projected nonlinear term
    vanishes at every substep and the step is exact: un -> F un."""
    g = O.Grid(32)
    u0 = abc_field(g)
    nu = 0.05
    r = O.Run(g, u0, nu, cfl=-1.0, kshift=1, dt_fixed=0.02)
    for _ in range(5):
        r.step_rk4()
    assert rel(r.uny, u0 * np.exp(-nu * r.time), g.mask()[None]) < 1e-12


def test_rk4_fourth_order_in_dt_inviscid(O):
    """With nu = 0 (F = 1) advanc_1..4 is the classical RK4 with a projection
    at each substep: the error falls 16-fold per halving of dt.  A wrong
    substep weight or a dropped term breaks the order."""
    g = O.Grid(32)
    u0 = tg_field(g)
    T = 0.4

    def run(dt):
        r = O.Run(g, u0, 0.0, cfl=-1.0, kshift=4, dt_fixed=dt)
        for _ in range(int(round(T / dt))):
            r.step_rk4()
        return r.uny

    ref = run(0.0125)
    e1 = np.linalg.norm((run(0.1) - ref).ravel())
    e2 = np.linalg.norm((run(0.05) - ref).ravel())
    assert 12.0 < e1 / e2 < 20.0


def test_rk4_first_order_in_dt_viscous(O):
    """With nu > 0 advanc_2 and advanc_3 add the nonlinear term to the
    unfactored un (no integrating factor after substep 1), so the scheme is
    first order in dt: the error halves with dt (a consistent
    integrating-factor RK4 would fall 16-fold)."""
    g = O.Grid(32)
    u0 = tg_field(g)
    T, nu = 0.4, 0.05

    def run(dt):
        r = O.Run(g, u0, nu, cfl=-1.0, kshift=4, dt_fixed=dt)
        for _ in range(int(round(T / dt))):
            r.step_rk4()
        return r.uny

    ref = run(0.00625)
    e = [np.linalg.norm((run(dt) - ref).ravel()) for dt in (0.05, 0.025)]
    assert 1.8 < e[0] / e[1] < 3.0


def test_rk4_scalar_pure_diffusion_is_exact(O):
    """u = 0, G = 0: the scalar terms vanish and RK4 (factor in substep 1
    only) gives each mode exp(-nu/pr k^2 dt) per step exactly."""
    n, nu, pr, dt = 32, 0.05, 0.7, 0.01
    g = O.Grid(n)
    rng = np.random.default_rng(9)
    phi0 = g.xktran(rng.standard_normal((1, n, n, n)), 1)
    phi0[:, ~g.mask()] = 0
    phi0[0, 0, 0] = 0
    ro = O.Run(g, g.zeros_cyl(3), nu, cfl=np.pi, kshift=4, dt_fixed=dt, phi=phi0, pr=[pr], grad=[(0, 0, 0)])
    for _ in range(3):
        ro.step_rk4()
    kx, ky, kz = kgrid(n, g.col_x, g.col_z)
    ref = phi0[0] * np.exp(-nu / pr * (kx * kx + ky * ky + kz * kz) * 3 * dt)
    assert rel(ro.phi_n[0], ref, g.mask()) < 1e-13


def sinusoidal_field(g):
    """pfield.f (hooks/initial_conditions): au = av = 1, aw = -2"""
    n = g.n
    x = np.arange(n) * 2 * np.pi / n
    Y, Z, X = np.meshgrid(x, x, x, indexing="ij")
    ux = np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), np.cos(X) * np.sin(Y) * np.cos(Z),
                   -2 * np.cos(X) * np.cos(Y) * np.sin(Z)])
    return g.xktran(ux, 3)


def test_reference_invariants_sinusoidal_ic(O):
    # README.spp:101-105: TKE must not increase; rms u = rms v (u<->v symmetry)
    g = O.Grid(32)
    u0 = sinusoidal_field(g)
    r = O.Run(g, u0, 0.01, cfl=np.pi, kshift=4)
    prev = g.spectrum(r.uny, 0.01)[2]
    tf = np.where(g.col_x == 0, 1.0, 2.0)[:, None]
    for _ in range(6):
        r.step()
        t = g.spectrum(r.uny, 0.01)[2]
        assert t <= prev * (1 + 1e-13)
        prev = t
        eu = np.sum(tf * np.abs(r.uny[0]) ** 2)
        ev = np.sum(tf * np.abs(r.uny[1]) ** 2)
        assert abs(eu - ev) < 1e-12 * eu


def test_fek_forcing_holds_low_shells(O):
    g = O.Grid(32)
    u0 = g.field(seed=11)
    r = O.Run(g, u0, 0.01, cfl=np.pi, kshift=1, kf_shell=3)
    target = r.ekinf.copy()
    for _ in range(3):
        r.step()
        ek = g.ek_shell(r.uny, 4)
        assert np.allclose(ek[1:3], target[1:3], rtol=1e-13)
        assert ek[0] == 0.0


def test_initial_field_spectrum_is_deterministic(O):
    # field.f:224-245: |u|^2 = ef^2 per mode, independent of the random phases
    g = O.Grid(32)
    e1 = g.spectrum(g.field(seed=1), 0.0)[0]
    e2 = g.spectrum(g.field(seed=2), 0.0)[0]
    assert np.allclose(e1, e2, rtol=1e-12, atol=1e-16)
    assert g.spectrum(g.field(seed=1), 0.0)[2] == pytest.approx(0.5, rel=1e-13)
    # Hermitian kx = 0 plane: the transform is real (Im part of kx=0 line vanishes)
    u = g.field(seed=3)
    back = g.xktran(g.kxtran(np.concatenate([u, g.zeros_cyl(2)]), 3), 3)
    assert rel(back, u, g.mask()[None]) < 1e-14


def test_vg_moments_closed_form(O):
    """vgmomt.f:68-143 moments of du/dx and dw/dz against the analytic
    derivatives of a few Fourier modes (kx >= 1) sampled on the grid."""
    n = 32
    g = O.Grid(n)
    # kx = 1, 1, -2 close a triad: du/dx has a nonzero third moment
    modes_u = [((1, 0, 0), 0.5), ((2, 0, 0), 0.3j), ((2, 1, -1), -0.25j)]
    modes_w = [((1, 0, 3), 0.5), ((2, -2, 1), 0.1 + 0.2j)]
    u = g.zeros_cyl(3)
    for comp, modes in ((0, modes_u), (2, modes_w)):
        for (kx, ky, kz), c in modes:
            a = next(i for i in range(g.num_al) if g.col_x[i] == kx and g.col_z[i] == kz % n)
            u[comp, a, ky % n] = c
    x = np.arange(n) * 2 * np.pi / n
    Y, Z, X = np.meshgrid(x, x, x, indexing="ij")

    def deriv(modes, axis):
        f = np.zeros((n, n, n))
        for k, c in modes:
            f += 2 * np.real(1j * k[axis] * c * np.exp(1j * (k[0] * X + k[1] * Y + k[2] * Z)))
        return f

    du, dw = g.vg_moments(u)
    for got, f in ((du, deriv(modes_u, 0)), (dw, deriv(modes_w, 2))):
        ref = np.array([np.mean(f ** p) for p in range(1, 9)])
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)
    # the third moment is not zero, so the sign of i k is pinned
    assert abs(du[2]) > 0.1


def test_scalar_shift_matches_the_velocity_shift(O):
    """the scalar phase factors equal the oracle's velocity shift tables
    (phshift_inplace of a field of ones on the mask)"""
    import oracle as orc
    n = 32
    g = O.Grid(n)
    gsh = np.array([0.3, 0.7, 0.1, 0.45, 0.05, 0.9])
    g.set_shifts(gsh)
    for k in (1, 2):
        w = g.zeros_cyl(5)
        w[:3] = 1.0
        g.phshift_inplace(w, 3, k)
        m = g.mask()
        assert np.abs((w[0] - orc.scalar_shift(g, gsh, k))[m]).max() < 1e-14


def test_scalar_pure_diffusion_is_exact(O):
    """u = 0, no mean gradient: the integrating factor makes every scalar
    mode decay as exp(-nu/pr k^2 t) exactly (proc4a/proc4b with zero terms)"""
    n, nu, pr, dt = 32, 0.05, 0.7, 0.01
    g = O.Grid(n)
    phi0 = random_cylinder(g, 1, 5)
    m = g.mask()
    phi0[0, 0, 0] = 0  # the mean is zeroed by proc4b
    phi0[0][~m] = 0
    ro = O.Run(g, g.zeros_cyl(3), nu, cfl=np.pi, kshift=4, dt_fixed=dt, phi=phi0, pr=[pr], grad=[(0, 0, 0)])
    for _ in range(3):
        ro.step()
    kx, ky, kz = kgrid(n, g.col_x, g.col_z)
    ref = phi0[0] * np.exp(-nu / pr * (kx * kx + ky * ky + kz * kz) * 3 * dt)
    assert rel(ro.phi_n[0], ref, m) < 1e-13


def test_scalar_terms_vs_numpy(O):
    """u.(G + grad phi) for band-limited fields (no aliasing) against a numpy
    evaluation on the grid, no phase shift"""
    import oracle as orc
    n = 32
    g = O.Grid(n)
    x = np.arange(n) * 2 * np.pi / n
    Y, Z, X = np.meshgrid(x, x, x, indexing="ij")
    vel = [np.cos(X + 2 * Y) + 0.3 * np.sin(Z), 0.5 * np.sin(2 * X - Z), np.cos(Y) - 0.2 * np.cos(X + Z)]
    phi = 0.7 * np.cos(X + Y + Z) + 0.4 * np.sin(3 * X - Z)
    dphi = [-0.7 * np.sin(X + Y + Z) + 1.2 * np.cos(3 * X - Z), -0.7 * np.sin(X + Y + Z),
            -0.7 * np.sin(X + Y + Z) - 0.4 * np.cos(3 * X - Z)]
    grad = (1.0, -0.5, 0.25)
    term = sum(vel[j] * (grad[j] + dphi[j]) for j in range(3))
    ux = g.zeros_phys(5)
    ux[:3] = vel
    phic = g.xktran(phi[None], 1)
    gsh = np.zeros(6)
    got = orc.scalar_terms(g, ux, phic, 1, [grad], gsh)[0]
    ref = g.xktran(term[None], 1)[0]
    # xktran itself is pinned against numpy rfftn (test_xktran_vs_numpy)
    assert rel(got, ref) < 1e-13
