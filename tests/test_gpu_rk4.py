"""RK4 (rkmethod = 4) through the C ABI against the FP64 oracle: four
substeps of stage_nonlinear + rk4_update (homogeneous/advanc_1.f ..
advanc_4.f, stages 3 and 4 with the shifts of stages 1 and 2) from the same
seeded field and shift offsets, with and without FEK forcing."""
import numpy as np
import pytest

from conftest import kgrid, rel

pytestmark = pytest.mark.gpu

TOL = {8: 1e-11, 4: 5e-6}


@pytest.mark.parametrize("n,prec,kf_shell,nsteps", [(32, 8, 0, 3), (64, 8, 3, 3), (64, 4, 3, 3), (128, 4, 0, 2)])
def test_rk4_steps_match_oracle(P, O, n, prec, kf_shell, nsteps):
    print("[gpu case] rk4", n, prec, kf_shell, nsteps, flush=True)
    g = O.Grid(n)
    u0 = g.field(seed=7)
    nu = 0.02
    ro = O.Run(g, u0, nu, cfl=np.pi, kshift=1, kf_shell=kf_shell, shift_seed=3)
    shifts = ro.shift_sequence(nsteps)
    ro = O.Run(g, u0, nu, cfl=np.pi, kshift=1, kf_shell=kf_shell, shift_seed=3)
    s = P.Solver(n, prec=prec, viscos=nu)
    s.put_state(u0)
    rg = P.Run4(s, nu, np.pi, kf_shell=kf_shell, ekinf=ro.ekinf)
    m = g.mask()[None]
    for k in range(nsteps):
        ro.step_rk4(shifts[k])
        rg.step(shifts[k])
        assert abs(rg.dt - ro.dt) <= 1e-6 * ro.dt
    assert rel(s.get_state(), ro.uny, m) < TOL[prec]
    # u = un after the step (advanc_4.f:149-163)
    assert rel(s.get_field(P.FIELD_U, 3), ro.uny, m) < TOL[prec]
    s.close()


def test_rk4_call_order(P, O):
    print("[gpu case] rk4 order", flush=True)
    g = O.Grid(32)
    s = P.Solver(32, prec=8)
    s.put_state(g.field(seed=1))
    s.set_shifts([0.0] * 6)
    with pytest.raises(P.PsdnsError, match="error -4"):
        s.rk4_update(1, 0.01, *s.step_factors(0.01))
    s.stage_nonlinear(1)
    with pytest.raises(P.PsdnsError, match="error -4"):
        s.rk4_update(2, 0.01)
    with pytest.raises(P.PsdnsError, match="error -1"):
        s.rk4_update(5, 0.01)
    s.close()


def scalar_ic(g, nc, seed):
    rng = np.random.default_rng(seed)
    phi = g.xktran(rng.standard_normal((nc, g.n, g.n, g.n)), nc)
    phi[:, ~g.mask()] = 0
    return phi


@pytest.mark.parametrize("n,prec,nsteps", [(32, 8, 2), (64, 4, 2)])
def test_rk4_scalars_match_oracle(P, O, n, prec, nsteps):
    """RK4 with two passive scalars (advanc_K's loops over i = 4 .. 3 + nc)."""
    print("[gpu case] rk4 scalars", n, prec, nsteps, flush=True)
    g = O.Grid(n)
    u0 = g.field(seed=31)
    phi0 = scalar_ic(g, 2, 7)
    pr, grad = [0.7, 1.3], [(1.0, 0.0, 0.0), (0.0, 0.5, -0.25)]
    nu = 0.02
    ro = O.Run(g, u0, nu, cfl=np.pi, kshift=1, shift_seed=3, phi=phi0, pr=pr, grad=grad)
    s = P.Solver(n, prec, viscos=nu, nc=2, pr=pr, grad=grad)
    s.put_state(np.concatenate([u0, phi0]))
    rg = P.Run4(s, nu, np.pi)
    m = g.mask()[None]
    for _ in range(nsteps):
        gsh = ro.next_shifts()
        ro.step_rk4(gsh)
        rg.step(gsh)
    st = s.get_state()
    ev, es = rel(st[:3], ro.uny, m), rel(st[3:], ro.phi_n, m)
    print(f"velocity rel L2 {ev:.2e}, scalars rel L2 {es:.2e}", flush=True)
    assert ev < TOL[prec] and es < TOL[prec]
    s.close()


def test_fortran_driver_rk4(P, O, tmp_path):
    """The Fortran driver with rk=4 (rkmethod = 4) through the ISO_C_BINDING
    module, against the oracle's RK4 from the same field and shifts."""
    import subprocess
    print("[gpu case] driver rk4", flush=True)
    if not P.DRIVER_PATH.exists():
        pytest.skip("Fortran driver not built")
    n, nsteps = 32, 3
    g = O.Grid(n)
    u0 = g.field(seed=5)
    ro = O.Run(g, u0, 0.01, cfl=np.pi, kshift=1, kf_shell=3, shift_seed=5)
    shifts = ro.shift_sequence(nsteps)
    ro = O.Run(g, u0, 0.01, cfl=np.pi, kshift=1, kf_shell=3, shift_seed=5)
    for k in range(nsteps):
        ro.step_rk4(shifts[k])
    np.ascontiguousarray(u0, np.complex128).tofile(tmp_path / "ic.bin")
    np.ascontiguousarray(shifts, np.float64).tofile(tmp_path / "shifts.bin")
    cmd = [str(P.DRIVER_PATH), f"n={n}", "prec=8", f"nsteps={nsteps}", "cfl=3.14159265358979", "viscos=0.01",
           "kshift=1", "kf_shell=3", "rk=4", f"ic={tmp_path / 'ic.bin'}", f"shifts={tmp_path / 'shifts.bin'}",
           f"out={tmp_path / 'out.bin'}", f"log={tmp_path / 'log'}"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    out = np.fromfile(tmp_path / "out.bin", np.complex128).reshape(3, g.num_al, n)
    assert rel(out, ro.uny, g.mask()[None]) < 1e-11


@pytest.mark.parametrize("n,nsteps", [(256, 2), (1024, 1)])
def test_rk4_size_independent_properties(P, n, nsteps):
    """FP32 RK4 steps at 256^3 and the full 1024^3 size: the state after
    advanc_4 is divergence-free, has zero mean and a Hermitian kx = kz = 0
    column, u = un, and the TKE stays inside the unforced decay budget."""
    print("[gpu case] rk4 properties", n, nsteps, flush=True)
    s = P.Solver(n, 4, viscos=0.01)
    s.init_field()
    cx, cz, _ = s.columns()
    kx, ky, kz = kgrid(n, cx, cz)
    rg = P.Run4(s, 0.01, np.pi)
    t0 = s.shell_stats(0)[2]
    rng = np.random.default_rng(1)
    for _ in range(nsteps):
        rg.step(rng.random(6))
    un = s.get_state()
    div = np.abs(kx * un[0] + ky * un[1] + kz * un[2])
    kmag = np.sqrt(kx * kx + ky * ky + kz * kz)
    scale = np.max(kmag * np.abs(un).max(axis=0))
    print(f"max |k.u| / max |k||u| = {div.max() / scale:.2e}", flush=True)
    assert div.max() < 1e-5 * scale
    del div, kmag
    assert np.abs(un[:, 0, 0]).max() == 0.0
    a0 = int(np.nonzero((cx == 0) & (cz == 0))[0][0])
    col = un[:, a0, :]
    mirror = np.conj(col[:, (n - np.arange(n)) % n])
    assert np.abs(col - mirror).max() < 1e-5 * np.abs(col).max()
    assert np.array_equal(s.get_field(P.FIELD_U, 3), un)
    _, _, t1, eps = s.shell_stats(n)
    assert 0 < t1 < t0
    assert t0 - t1 < 1.2 * eps * sum(rg.dts) + 1e-3 * t0
    s.close()
