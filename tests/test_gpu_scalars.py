"""GPU parity of the passive-scalar path (nc > 0) against the FP64 oracle:
the scalar advection term u.(G + grad phi) (itransform.f:190-203,522-531),
its wavespace factor (wavespace_module.f:347-352) and the scalar updates
(proc4a.f:506-520, proc4b.f:69-74,183-205) beside the velocity step.
Scalar fields are made Hermitian (the forward transform of a real field) and
compared on the mask, where the reference keeps them."""
import sys

import numpy as np
import pytest

from conftest import kgrid, rel

pytestmark = pytest.mark.gpu

TOL = {8: 1e-11, 4: 5e-6}


def banner(*a):
    print("[gpu case]", *a, flush=True)
    sys.stdout.flush()


def scalar_ic(g, nc, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((nc, g.n, g.n, g.n))
    phi = g.xktran(f, nc)
    phi[:, ~g.mask()] = 0
    return phi


# 128^3 and 256^3: work-groups of several waves in the scalar x pass and
# the gradient y pass (at 32^3 and 64^3 each is one partial wave); 256^3
# FP32 over 10 steps is the two-scalar configuration's parity case
@pytest.mark.parametrize("n,prec,nsteps", [(32, 8, 1), (64, 8, 3), (64, 4, 3), (128, 8, 2), (256, 4, 10)])
def test_scalar_steps_vs_oracle(O, P, n, prec, nsteps):
    banner("scalars nc=2", n, prec, nsteps)
    g = O.Grid(n)
    u0 = g.field(seed=31)
    phi0 = scalar_ic(g, 2, 7)
    pr, grad = [0.7, 1.3], [(1.0, 0.0, 0.0), (0.0, 0.5, -0.25)]
    viscos = 0.02
    ro = O.Run(g, u0, viscos, cfl=np.pi, kshift=1, shift_seed=3, phi=phi0, pr=pr, grad=grad)
    s = P.Solver(n, prec, viscos=viscos, nc=2, pr=pr, grad=grad)
    try:
        s.put_state(np.concatenate([u0, phi0]))
        rg = P.Run(s, viscos, np.pi)
        m = g.mask()[None]
        for _ in range(nsteps):
            gsh = ro.next_shifts()
            ro.step(gsh)
            rg.step(gsh)
        st = s.get_state()
        ev, es = rel(st[:3], ro.uny, m), rel(st[3:], ro.phi_n, m)
        print(f"velocity rel L2 {ev:.2e}, scalars rel L2 {es:.2e}", flush=True)
        assert ev < TOL[prec] and es < TOL[prec]
        assert np.allclose(rg.dts, ro.dts, rtol=1e-12 if prec == 8 else 1e-6)
    finally:
        s.close()


def test_scalar_pure_diffusion_on_gpu(O, P):
    """u = 0, G = 0: each scalar mode decays as exp(-nu/pr k^2 t) exactly"""
    banner("scalar diffusion")
    n, nu, pr, dt = 32, 0.05, 0.7, 0.01
    g = O.Grid(n)
    phi0 = scalar_ic(g, 1, 9)
    phi0[0, 0, 0] = 0
    s = P.Solver(n, 8, viscos=nu, nc=1, pr=[pr], grad=[(0, 0, 0)])
    try:
        s.put_state(np.concatenate([g.zeros_cyl(3), phi0]))
        xf, yf, zf = s.step_factors(dt)
        for _ in range(3):
            s.set_shifts(np.zeros(6))
            s.stage_nonlinear(1)
            s.stage_update(1, dt, xf, yf, zf)
            s.stage_nonlinear(2)
            s.stage_update(2, dt, xf, yf, zf)
        kx, ky, kz = kgrid(n, g.col_x, g.col_z)
        ref = phi0[0] * np.exp(-nu / pr * (kx * kx + ky * ky + kz * kz) * 3 * dt)
        assert rel(s.get_state()[3], ref, g.mask()) < 1e-13
    finally:
        s.close()


def test_scalars_1024_fp32_properties(P):
    """One RK2 step at the full 1024^3 FP32 size with two scalars (Sc = 1,
    mean gradient (1, 0, 0), zero initial fluctuation): the scalars keep
    zero mean and a Hermitian kx = kz = 0 column, and to first order in dt
    each equals -dt (u_x^n + u_x^(n+1)) / 2 (the mean-gradient source
    integrated by the RK2 step; advection of the scalar and diffusion are
    O(dt |grad u|) corrections)."""
    banner("1024^3 fp32 two scalars, one step")
    n = 1024
    s = P.Solver(n, 4, viscos=3.2e-5, nc=2, pr=[1.0, 1.0], grad=[(1.0, 0.0, 0.0), (1.0, 0.0, 0.0)])
    try:
        s.init_field()
        ux0 = s.get_state()[0].copy()
        rg = P.Run(s, 3.2e-5, np.pi)
        rg.step(np.random.default_rng(1).random(6))
        st = s.get_state()
        cx, cz, _ = s.columns()
        a0 = int(np.nonzero((cx == 0) & (cz == 0))[0][0])
        dt = rg.dts[-1]
        pred = -0.5 * dt * (ux0 + st[0])
        del ux0
        for i in (3, 4):
            phi = st[i]
            assert np.abs(phi[0, 0]) == 0.0
            col = phi[a0, :]
            mirror = np.conj(col[(n - np.arange(n)) % n])
            assert np.abs(col - mirror).max() < 1e-5 * np.abs(col).max()
            err = np.linalg.norm(phi - pred) / np.linalg.norm(pred)
            print(f"scalar {i - 2}: |phi + dt (u_x^n + u_x^n+1)/2| / |.| = {err:.3e} (dt {dt:.3e})", flush=True)
            assert err < 0.1
        assert np.array_equal(st[3], st[4])  # same Sc and gradient: same bits
    finally:
        s.close()
