"""Oracle pins (CPU) for the pieces the transform/RK2 checks leave open:

- shell binning of the FP64 statistics and the FEK shells
  (PSDNS/ek_shell.f:66,76, PSDNS/homogeneous/sptr.f:127-131,164): single
  Fourier modes at chosen (kx, ky, kz) land in a hand-tabulated shell with
  the tfact weight of their kx (1 at kx = 0, 2 otherwise, comsp_set.f:85-87);
- tfact again through Parseval: the shell energies of a random field sum to
  half the grid mean of u^2 + v^2 + w^2 of its inverse transform (kxtran is
  pinned against numpy irfftn);
- the passive-scalar update's -(dt/2) conj(sx sy sz)
  (PSDNS/homogeneous/wavespace_module.f:347-352) with the phase shifts on:
  a shear flow u = (A sin y, 0, 0) with mean gradient G = (1, 0, 0) and a
  zero initial scalar has the closed form
      phi = -A sin y (exp(-nu t) - exp(-kappa t)) / (kappa - nu),
  which IF-RK2 reproduces to second order in dt.
"""
import numpy as np
import pytest

from conftest import kgrid, random_cylinder, rel

# (kx, ky, kz) -> 1-based shell: shell ik holds |k| in [ik - 1.5, ik - 0.5)
SHELLS = {
    (0, 0, 0): 1,
    (1, 0, 0): 2,
    (0, 1, 1): 2,   # |k| = 1.414
    (1, 1, 1): 3,   # 1.732
    (0, 0, 2): 3,   # 2.000
    (0, 2, 1): 3,   # 2.236, kx = 0
    (1, 2, 1): 3,   # 2.449
    (2, 2, 0): 4,   # 2.828
    (0, 0, 3): 4,   # 3.000
    (1, 3, 2): 5,   # 3.742
    (3, 3, 3): 6,   # 5.196
    (4, 4, 2): 7,   # 6.000
}


def column_of(g, kx, kz):
    z = kz % g.n
    hits = np.nonzero((g.col_x == kx) & (g.col_z == z))[0]
    assert len(hits) == 1
    return int(hits[0])


def single_mode(g, kx, ky, kz, comp, c):
    u = g.zeros_cyl(3)
    u[comp, column_of(g, kx, kz), ky % g.n] = c
    return u


@pytest.mark.parametrize("mode,shell", sorted(SHELLS.items()))
def test_single_mode_shell_and_weight(O, mode, shell):
    g = O.Grid(32)
    kx, ky, kz = mode
    c = 0.3 - 0.7j
    nu = 0.02
    u = single_mode(g, kx, ky, kz, 1, c)
    tfact = 1.0 if kx == 0 else 2.0
    e = tfact * abs(c) ** 2 / 2.0
    k2 = kx * kx + ky * ky + kz * kz
    ek, dk, tke, eps = g.spectrum(u, nu, 16)
    want = np.zeros(16)
    want[shell - 1] = e
    assert np.array_equal(np.nonzero(ek)[0], np.nonzero(want)[0])
    assert ek[shell - 1] == pytest.approx(e, rel=1e-15)
    assert dk[shell - 1] == pytest.approx(2 * nu * k2 * e, rel=1e-15, abs=1e-300)
    assert tke == pytest.approx(e, rel=1e-15)
    assert eps == pytest.approx(2 * nu * k2 * e, rel=1e-15, abs=1e-300)
    # the FEK shell sums use the same bins and weight (ek_shell.f:131-168)
    fek = g.ek_shell(u, 8)
    want8 = np.zeros(8)
    if shell <= 8:
        want8[shell - 1] = e
    assert np.allclose(fek, want8, rtol=1e-15, atol=0)


def test_shell_weight_through_parseval(O):
    """sum_k tfact |u_k|^2 / 2 over the half spectrum = <u.u>/2 on the grid"""
    g = O.Grid(32)
    u = random_cylinder(g, 3, 21)
    _, _, tke, _ = g.spectrum(u, 0.0, 1)
    phys = g.kxtran(np.concatenate([u, g.zeros_cyl(2)]), 3)
    assert tke == pytest.approx(0.5 * np.mean(np.sum(phys ** 2, axis=0)), rel=1e-13)
    ek, _, _, _ = g.spectrum(u, 0.0, g.n)
    assert ek.sum() == pytest.approx(tke, rel=1e-13)


def shear_field(g, A):
    """u = (A sin y, 0, 0) on the cylinder: kx = kz = 0, ky = +-1"""
    u = g.zeros_cyl(3)
    a = column_of(g, 0, 0)
    u[0, a, 1] = A / 2j
    u[0, a, g.n - 1] = -A / 2j
    return u


def shear_scalar_run(O, g, dt, T, gsh_seq, nu, pr, A):
    ro = O.Run(g, shear_field(g, A), nu, cfl=-1.0, kshift=4, dt_fixed=dt,
               phi=g.zeros_cyl(1), pr=[pr], grad=[(1.0, 0.0, 0.0)])
    nsteps = int(round(T / dt))
    for i in range(nsteps):
        ro.step(gsh_seq[i % len(gsh_seq)])
    return ro


def test_scalar_shear_flow_closed_form_with_shifts(O):
    n, nu, pr, A, T = 16, 0.1, 0.5, 1.0, 0.4
    kappa = nu / pr
    g = O.Grid(n)
    rng = np.random.default_rng(3)
    # shifts along y matter here (the field varies along y only); all three on
    gsh_seq = [rng.random(6) for _ in range(8)]
    a = column_of(g, 0, 0)
    m = g.mask()
    exact = -A * (np.exp(-nu * T) - np.exp(-kappa * T)) / (kappa - nu)
    errs = []
    for dt in (0.04, 0.02):
        ro = shear_scalar_run(O, g, dt, T, gsh_seq, nu, pr, A)
        # the velocity is an exact Stokes solution (nonlinear term zero)
        assert rel(ro.uny, shear_field(g, A * np.exp(-nu * T)), m[None]) < 1e-13
        want = g.zeros_cyl(1)[0]
        want[a, 1] = exact / 2j
        want[a, n - 1] = -exact / 2j
        errs.append(rel(ro.phi_n[0], want, m))
        # nothing leaks into other modes
        rest = ro.phi_n[0].copy()
        rest[a, 1] = rest[a, n - 1] = 0
        assert np.abs(rest).max() < 1e-14
    assert errs[0] < 2e-3
    assert 3.3 < errs[0] / errs[1] < 4.7


# ---------------------------------------------------------------------------
# EPFOR forcing increments (proc4a.f:588-640, proc4b.f:118-163, advanc_2/3)
# ---------------------------------------------------------------------------
def fwave(i, kfor):
    """1-based index of for(x, iy, iz, i) along y or z -> wavenumber"""
    return i - 1 if i <= kfor else i - 2 * kfor


def encoded_forcing(kfor):
    """for(x, iy, iz, i) = kx + ky/64 + 1j (kz + 10) (i + 1)/1024 (exact in
    binary): every entry names its own wavenumbers and component"""
    k2fo = 2 * kfor - 1
    f = np.zeros((3, k2fo, k2fo, kfor), complex)
    for i in range(3):
        for iz in range(1, k2fo + 1):
            for iy in range(1, k2fo + 1):
                for x in range(1, kfor + 1):
                    f[i, iz - 1, iy - 1, x - 1] = enc(x - 1, fwave(iy, kfor), fwave(iz, kfor), i)
    return f


def enc(kx, ky, kz, i):
    return kx + ky * 0.015625 + 1j * ((kz + 10) * (i + 1) * 0.0009765625)


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_epfor_index_map(O, mode):
    n, kfor, kforce = 32, 4, 2.6
    g = O.Grid(n)
    uy = g.zeros_cyl(5)
    uny = g.zeros_cyl(3)
    g.epfor(uy, uny, kfor, kforce, encoded_forcing(kfor), mode)
    kx, ky, kz = kgrid(n, g.col_x, g.col_z)
    kx, ky, kz = (np.broadcast_to(k, (g.num_al, n)) for k in (kx, ky, kz))
    box = (kx < kfor) & (np.abs(ky) < kfor) & (np.abs(kz) < kfor) & g.mask()
    if mode == 1:
        box &= kx ** 2 + ky ** 2 + kz ** 2 <= kforce ** 2
    for i in range(3):
        want = np.where(box, enc(kx, ky, kz, i), 0)
        if mode == 2:
            want[0, 0] = 0  # zero mean after the add (proc4b.f:200-205)
        assert np.array_equal(uy[i], want)
        assert np.array_equal(uny[i], want if mode == 2 else 0 * want)


def test_epfor_tally_is_the_energy_input(O):
    """0.5 sum_i efkz(k, i) = E_k(u + f) - E_k(u) - E_k(f) per shell, with
    E_k the (pinned) shell spectrum: the tally is sum tfact Re(u conj f)"""
    import oracle as orc
    n, kfor, kforce = 32, 5, 4.2
    g = O.Grid(n)
    rng = np.random.default_rng(4)
    forc = orc.epfor_increments(kfor, kforce, rng, amp=0.3)
    u = g.zeros_cyl(5)
    u[:3] = random_cylinder(g, 3, 17)
    before = g.spectrum(np.ascontiguousarray(u[:3]), 0.0, n)[0]
    fonly = g.zeros_cyl(5)
    g.epfor(fonly, g.zeros_cyl(3), kfor, kforce, forc, 3)
    ef = g.spectrum(np.ascontiguousarray(fonly[:3]), 0.0, n)[0]
    efkz = g.epfor(u, g.zeros_cyl(3), kfor, kforce, forc, 1)
    after = g.spectrum(np.ascontiguousarray(u[:3]), 0.0, n)[0]
    tally = 0.5 * efkz.sum(axis=0)
    scale = np.abs(before).max()
    assert np.abs(tally[:n // 2] - (after - before - ef)[:n // 2]).max() < 1e-12 * scale
    assert np.abs(tally).max() > 1e-6 * scale  # far above the tolerance: not vacuous
    assert np.all(tally[int(kforce + 1.5):] == 0)


def test_epfor_increments_are_solenoidal_and_hermitian(O):
    import oracle as orc
    n, kfor, kforce = 32, 5, 4.2
    g = O.Grid(n)
    forc = orc.epfor_increments(kfor, kforce, np.random.default_rng(9))
    f = g.zeros_cyl(5)
    g.epfor(f, g.zeros_cyl(3), kfor, kforce, forc, 3)
    kx, ky, kz = kgrid(n, g.col_x, g.col_z)
    assert np.abs(kx * f[0] + ky * f[1] + kz * f[2]).max() < 1e-13 * np.abs(f).max()
    # conjugate symmetry on kx = 0 (force.f:140-156): the field is real
    back = g.xktran(g.kxtran(f, 3), 3)
    assert rel(back, f[:3]) < 1e-14
    assert np.abs(f[:, 0, 0]).max() == 0  # k = 0 is never forced


@pytest.mark.parametrize("n", [64, 256, 1024])
def test_fp32_shell_index_equals_fp64(n):
    """The GPU shell statistics form ik = int(|k| + 1.5) (PSDNS/ek_shell.f:66)
    with a correctly rounded FP32 square root of the integer |k|^2: for every
    |k|^2 a mask mode can reach (9 |k|^2 <= 2 n^2) it equals the FP64 index,
    because sqrt(k2) + 1.5 stays at least 2.5e-4 away from an integer."""
    k2 = np.arange(0, 2 * n * n // 9 + 1, dtype=np.int64)
    fp64 = np.floor(np.sqrt(k2.astype(np.float64)) + 1.5).astype(np.int64)
    fp32 = np.floor(np.sqrt(k2.astype(np.float32)) + np.float32(1.5)).astype(np.int64)
    assert np.array_equal(fp32, fp64)
    t = np.sqrt(k2.astype(np.float64)) + 1.5
    assert np.min(np.abs(t - np.round(t))) > 2.5e-4
