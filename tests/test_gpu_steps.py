"""Multi-step GPU runs vs the FP64 oracle from the same seeded field, the
reference's own run-time invariants, size-independent properties at 256^3,
and the Fortran ISO_C_BINDING driver as a drop-in.

Tolerances (relative L2 of the spectral velocity over the mask):
  FP64 GPU vs FP64 oracle: 1e-11 after 20 steps at 64^3;
  FP32 GPU vs FP64 oracle: 1e-5 after 10 steps (the stated target).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import REPO, kgrid, rel

pytestmark = pytest.mark.gpu


def banner(*a):
    print("[gpu case]", *a, flush=True)
    sys.stdout.flush()


def run_pair(O, P, n, prec, nsteps, kshift=1, kf_shell=3, seed=122929298, viscos=0.01):
    g = O.Grid(n)
    u0 = g.field(seed=seed)
    ro = O.Run(g, u0, viscos, cfl=np.pi, kshift=kshift, kf_shell=kf_shell, shift_seed=seed)
    s = P.Solver(n, prec, viscos=viscos)
    s.put_state(u0)
    rg = P.Run(s, viscos, np.pi, kf_shell=kf_shell, ekinf=ro.ekinf)
    errs = []
    m = g.mask()[None]
    for _ in range(nsteps):
        gsh = ro.next_shifts()
        ro.step(gsh)
        rg.step(gsh)
        errs.append(rel(s.get_state(), ro.uny, m))
    return g, s, ro, rg, errs


def test_fp64_20_steps_64(O, P):
    banner("fp64 20 steps 64^3")
    g, s, ro, rg, errs = run_pair(O, P, 64, 8, 20)
    print("rel L2 per step:", ["%.2e" % e for e in errs])
    assert max(errs) < 1e-11
    assert np.allclose(rg.dts, ro.dts, rtol=1e-12)
    ek, dk, tke, eps = s.shell_stats(64)
    rek, rdk, rtke, reps = g.spectrum(ro.uny, 0.01)
    assert tke == pytest.approx(rtke, rel=1e-12)
    assert eps == pytest.approx(reps, rel=1e-12)
    s.close()


@pytest.mark.parametrize("n", [64, 128])
def test_fp32_10_steps_vs_fp64_oracle(O, P, n):
    banner("fp32 10 steps", n)
    g, s, ro, rg, errs = run_pair(O, P, n, 4, 10)
    print("rel L2 per step:", ["%.2e" % e for e in errs])
    assert errs[-1] < 1e-5
    ek, dk, tke, eps = s.shell_stats(n)
    rek, rdk, rtke, reps = g.spectrum(ro.uny, 0.01)
    assert tke == pytest.approx(rtke, rel=1e-5)
    assert eps == pytest.approx(reps, rel=1e-5)
    sel = rek > 1e-8 * rek.max()
    assert np.max(np.abs(ek[sel] / rek[sel] - 1)) < 1e-4
    s.close()


def skewness(g, uy, cx, cz):
    """mean longitudinal velocity-derivative skewness <(du_i/dx_i)^3> /
    <(du_i/dx_i)^2>^(3/2) over i = x, y, z, from the spectral field (oracle
    inverse transform in FP64)"""
    kx, ky, kz = kgrid(g.n, cx, cz)
    out = []
    for i, k in enumerate((kx, ky, kz)):
        d = np.zeros((1,) + uy.shape[1:], np.complex128)
        d[0] = 1j * k * uy[i]
        f = g.kxtran(d, 1)[0]
        out.append(np.mean(f ** 3) / np.mean(f ** 2) ** 1.5)
    return float(np.mean(out))


@pytest.mark.parametrize("n,prec,nsteps", [(256, 4, 10), (256, 8, 10), (512, 4, 10), (512, 8, 10)])
def test_forced_steps_vs_fp64_oracle(O, P, n, prec, nsteps):
    """Forced isotropic turbulence at 256^3 and 512^3, viscosity from
    kmax eta = 1.5, FEK forcing of shells k <= 2, RK2 steps with phase
    shifts: GPU (FP32 or FP64) vs the FP64 oracle from the same seeded field.
    Reports the spectral relative L2 error, E(k), dissipation and the
    velocity-derivative skewness (and, on the GPU, the du/dx and dw/dz
    moments against the oracle's)."""
    banner(f"forced {n}^3 fp{8 * prec} {nsteps} steps")
    g = O.Grid(n)
    u0 = g.field(seed=122929298)
    _, _, _, eps1 = g.spectrum(u0, 1.0)
    viscos = float(np.sqrt(eps1) * (1.5 / (np.sqrt(2.0) / 3.0 * n)) ** 2)
    ro = O.Run(g, u0, viscos, cfl=np.pi, kshift=1, kf_shell=3, shift_seed=11)
    s = P.Solver(n, prec, viscos=viscos)
    s.put_state(u0)
    rg = P.Run(s, viscos, np.pi, kf_shell=3, ekinf=ro.ekinf)
    m = g.mask()[None]
    errs = []
    for _ in range(nsteps):
        gsh = ro.next_shifts()
        ro.step(gsh)
        rg.step(gsh)
        errs.append(rel(s.get_state(), ro.uny, m))
    ek, dk, tke, eps = s.shell_stats(n)
    rek, rdk, rtke, reps = g.spectrum(ro.uny, viscos)
    sel = rek > 1e-10 * rek.max()
    ekerr = float(np.max(np.abs(ek[sel] / rek[sel] - 1)))
    cx, cz, _ = s.columns()
    if n <= 256:  # (the numpy evaluation is slow at 512^3; the moments below cover it)
        sk, rsk = skewness(g, s.get_state().astype(np.complex128), cx, cz), skewness(g, ro.uny, cx, cz)
    else:
        sk = rsk = 0.0
    print(f"fp{8 * prec}: rel L2 per step {['%.2e' % e for e in errs]}; E(k) max rel {ekerr:.2e}; "
          f"eps {eps:.6e} vs {reps:.6e}; skewness {sk:.6f} vs {rsk:.6f}", flush=True)
    tol = 1e-5 if prec == 4 else 1e-11
    assert errs[-1] < tol
    assert np.allclose(rg.dts, ro.dts, rtol=1e-6 if prec == 4 else 1e-12)
    assert ekerr < (1e-4 if prec == 4 else 1e-10)
    assert eps == pytest.approx(reps, rel=tol)
    assert tke == pytest.approx(rtke, rel=tol)
    assert abs(sk - rsk) < (1e-4 if prec == 4 else 1e-10)
    for got, ref in zip(s.vg_moments(), g.vg_moments(ro.uny)):
        skg, skr = got[2] / got[1] ** 1.5, ref[2] / ref[1] ** 1.5
        assert abs(skg - skr) < (1e-4 if prec == 4 else 1e-10)
        assert got[3] / got[1] ** 2 == pytest.approx(ref[3] / ref[1] ** 2, rel=1e-4 if prec == 4 else 1e-10)
    s.close()


def test_fek_forcing_on_gpu(O, P):
    banner("fek forcing 64^3 fp64")
    n = 64
    s = P.Solver(n, 8, viscos=0.01)
    s.init_field()
    target = s.ek_shell(4)
    rg = P.Run(s, 0.01, np.pi, kf_shell=3, ekinf=target)
    for _ in range(3):
        rg.step(np.array([0.1, 0.2, 0.3, 0.6, 0.7, 0.8]))
        ek = s.ek_shell(4)
        assert np.allclose(ek[1:3], target[1:3], rtol=1e-12)
        assert ek[0] == 0.0
    s.close()


def test_reference_invariants_sinusoidal_ic(O, P):
    # README.spp:101-105 on the GPU: TKE does not increase, rms u = rms v
    banner("sinusoidal IC invariants 64^3 fp64")
    n = 64
    g = O.Grid(n)
    x = np.arange(n) * 2 * np.pi / n
    Y, Z, X = np.meshgrid(x, x, x, indexing="ij")
    ux = np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), np.cos(X) * np.sin(Y) * np.cos(Z),
                   -2 * np.cos(X) * np.cos(Y) * np.sin(Z)])
    u0 = g.xktran(ux, 3)
    s = P.Solver(n, 8, viscos=0.01)
    s.put_state(u0)
    rg = P.Run(s, 0.01, np.pi)
    prev = s.shell_stats(0)[2]
    tf = np.where(g.col_x == 0, 1.0, 2.0)[:, None]
    for _ in range(6):
        rg.step(np.zeros(6))
        t = s.shell_stats(0)[2]
        assert t <= prev * (1 + 1e-13)
        prev = t
        un = s.get_state()
        eu, ev = np.sum(tf * np.abs(un[0]) ** 2), np.sum(tf * np.abs(un[1]) ** 2)
        assert abs(eu - ev) < 1e-12 * eu
    s.close()


def test_stokes_decay_beltrami_on_gpu(O, P):
    banner("beltrami decay 64^3 fp64/fp32")
    n, nu = 64, 0.05
    g = O.Grid(n)
    x = np.arange(n) * 2 * np.pi / n
    Y, Z, X = np.meshgrid(x, x, x, indexing="ij")
    ux = np.stack([np.sin(Z) + 0.4 * np.cos(Y), 0.7 * np.sin(X) + np.cos(Z), 0.4 * np.sin(Y) + 0.7 * np.cos(X)])
    u0 = g.xktran(ux, 3)
    for prec, tol in ((8, 1e-12), (4, 2e-6)):
        s = P.Solver(n, prec, viscos=nu)
        s.put_state(u0)
        t = 0.0
        for _ in range(5):
            s.set_shifts(np.array([0.3, 0.1, 0.7, 0.8, 0.6, 0.2]))
            dt = 0.02
            xf, yf, zf = s.step_factors(dt)
            s.stage_nonlinear(1)
            s.stage_update(1, dt, xf, yf, zf)
            s.stage_nonlinear(2)
            s.stage_update(2, dt, xf, yf, zf)
            t += dt
        assert rel(s.get_state(), u0 * np.exp(-nu * t), g.mask()[None]) < tol
        s.close()


@pytest.mark.parametrize("n,nsteps", [(256, 3), (1024, 1)])
def test_size_independent_properties(O, P, n, nsteps):
    """FP32 RK2 steps at 256^3 and at the full 1024^3 size: divergence-free
    state, zero mean, Hermitian kx = 0, kz = 0 column (real physical field),
    TKE within the unforced decay budget; at 1024^3 also the kxtran/xktran
    round trip of one variable (identity on the mask)."""
    banner(f"{n}^3 fp32 properties, {nsteps} steps")
    s = P.Solver(n, 4, viscos=0.01)
    s.init_field()
    cx, cz, _ = s.columns()
    kx, ky, kz = kgrid(n, cx, cz)
    if n >= 1024:
        u0 = s.get_state()[0].copy()
        s.put_field(P.FIELD_U, u0[None])
        s.kxtran(1)
        s.xktran(1)
        back = s.get_field(P.FIELD_U, 1)[0]
        m = (9 * (kx * kx + ky * ky + kz * kz) <= 2 * n * n) & (ky != -n // 2) & (kz != -n // 2)
        err = np.linalg.norm((back - u0)[m]) / np.linalg.norm(u0[m])
        print(f"kxtran/xktran round trip at {n}^3: rel L2 {err:.2e}", flush=True)
        assert err < 1e-6
        del u0, back, m
        s.put_state(s.get_state())  # u = un again for the step
    rg = P.Run(s, 0.01, np.pi)
    t0 = s.shell_stats(0)[2]
    rng = np.random.default_rng(0)
    for _ in range(nsteps):
        rg.step(rng.random(6))
    un = s.get_state()  # complex64
    div = np.abs(kx * un[0] + ky * un[1] + kz * un[2])
    kmag = np.sqrt(kx * kx + ky * ky + kz * kz)
    scale = np.max(kmag * np.abs(un).max(axis=0))
    print(f"max |k.u| / max |k||u| = {div.max() / scale:.2e}", flush=True)
    assert div.max() < 1e-5 * scale
    del div, kmag
    assert np.abs(un[:, 0, 0]).max() == 0.0
    # kx = 0, kz = 0 column: u(-ky) = conj u(ky) up to FP32 round-off
    a0 = int(np.nonzero((cx == 0) & (cz == 0))[0][0])
    col = un[:, a0, :]
    mirror = np.conj(col[:, (n - np.arange(n)) % n])
    assert np.abs(col - mirror).max() < 1e-5 * np.abs(col).max()
    ek, dk, t1, eps = s.shell_stats(n)
    assert 0 < t1 < t0
    assert t0 - t1 < 1.2 * eps * sum(rg.dts) + 1e-3 * t0
    s.close()


def test_fortran_driver_matches_oracle(O, P, tmp_path):
    banner("fortran driver 64^3 fp64 5 steps")
    n, nsteps = 64, 5
    g = O.Grid(n)
    u0 = g.field(seed=5)
    ro = O.Run(g, u0, 0.01, cfl=np.pi, kshift=1, kf_shell=3, shift_seed=5)
    shifts = ro.shift_sequence(nsteps)
    ro2 = O.Run(g, u0, 0.01, cfl=np.pi, kshift=1, kf_shell=3, shift_seed=5)
    for k in range(nsteps):
        ro2.step(shifts[k])
    # Fortran column-major uy(ny, num_al, 3) == C-order (3, num_al, ny)
    np.ascontiguousarray(u0, np.complex128).tofile(tmp_path / "ic.bin")
    np.ascontiguousarray(shifts, np.float64).tofile(tmp_path / "shifts.bin")
    cmd = [str(P.DRIVER_PATH), f"n={n}", "prec=8", f"nsteps={nsteps}", "cfl=3.14159265358979",
           "viscos=0.01", "kshift=1", "kf_shell=3", f"ic={tmp_path / 'ic.bin'}",
           f"shifts={tmp_path / 'shifts.bin'}", f"out={tmp_path / 'out.bin'}", f"log={tmp_path / 'log'}",
           f"outpen={tmp_path / 'outpenc.p0'}"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    out = np.fromfile(tmp_path / "out.bin", np.complex128).reshape(3, g.num_al, n)
    assert rel(out, ro2.uny, g.mask()[None]) < 1e-11
    # the checkpoint the driver wrote at the end (main.f:604-608) holds the same state
    sys.path.insert(0, str(REPO / "oracle"))
    import pencil
    chk, fprec = pencil.read_inpen(tmp_path / "outpenc.p0", n, g.col_x, g.col_z, 3)
    assert fprec == 8 and np.array_equal(chk[:, pencil.held(n, g.col_z)], out[:, pencil.held(n, g.col_z)])
    log = (tmp_path / "log").read_text().splitlines()
    assert len([l for l in log if l.startswith("cpu at step")]) == nsteps


# FP64 at 64^3 and 128^3: the fused step has been seen to differ between
# reruns in the last, partly stored kx block of the x pass's second product
# (differences of about 1e-27 on values of about 1e-11, far below every
# parity tolerance); the cause is not found.  Non-strict: the variation
# does not show on every run.
@pytest.mark.parametrize("n,prec", [(32, 8), (64, 8), (128, 8), (256, 8), (64, 4), (256, 4)])
def test_fused_step_reproducible(P, O, n, prec):
    """The same state stepped three times on one context gives the same bits.
    The x pass stages each variable into LDS with asynchronous copies that
    several waves read, so every wave waits for its own copies before the
    publishing barrier; a missing wait shows up here as run-to-run
    differences in the highest kx modes.  In FP64 the forward x transform
    stores 16 bytes per lane: with the per-j offset in the store's scalar
    offset register, the low dword of a stored value was sometimes the one
    the next instruction wrote (64^3 and 128^3, last kx block of X slot 1);
    the offset now goes into the vector offset."""
    print("[gpu case] reproducible", n, prec, flush=True)
    g = O.Grid(n)
    u0 = g.field(seed=9)
    gsh = [0.1, 0.2, 0.3, 0.6, 0.7, 0.8]
    s = P.Solver(n, prec=prec, viscos=0.01)
    out = []
    for _ in range(3):
        s.put_state(u0)
        P.Run(s, 0.01, np.pi).step(gsh)
        out.append(s.get_state())
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[1], out[2])
    s.close()


@pytest.mark.parametrize("n,prec", [(32, 8), (64, 8), (128, 8), (256, 8), (64, 4), (256, 4)])
def test_forced_step_reproducible(P, O, n, prec):
    """Two forced steps (FEK shells k <= 2) from the same state give the
    same bits: the shell energies the rescale uses are fixed-order FP64
    sums and velmax is an order-free maximum, so no reduction depends on
    scheduling"""
    print("[gpu case] forced reproducible", n, prec, flush=True)
    g = O.Grid(n)
    u0 = g.field(seed=12)
    s = P.Solver(n, prec=prec, viscos=0.01)
    try:
        s.put_state(u0)
        ekinf = s.ek_shell(4)
        out = []
        for _ in range(2):
            s.put_state(u0)
            r = P.Run(s, 0.01, np.pi, kf_shell=3, ekinf=ekinf)
            rng = np.random.default_rng(3)
            for _ in range(2):
                r.step(rng.random(6))
            out.append((s.get_state(), s.shell_stats(n)))
        assert np.array_equal(out[0][0], out[1][0])
        assert np.array_equal(out[0][1][0], out[1][1][0]) and out[0][1][3] == out[1][1][3]
    finally:
        s.close()
