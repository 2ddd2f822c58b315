"""Lagrangian particles (hooks/LAG_LOCAL) through the C ABI against the FP64
oracle's spline path (solsp1/splx/splz/sply, spcal_local, rk_advpos):
the GPU forms the same periodic-spline coefficients by the exact spectral
inverse of the (1,4,1) systems, so the interpolated velocities agree to
round-off; the reference's RK2 particle step runs with the Eulerian RK2
steps (partsp.f:217-299).  Config 4's shape: 512^3 FP32 with 2^20
particles uniform in the box, within 1e-6 relative."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,prec,nsteps", [(32, 8, 3), (64, 8, 3), (128, 4, 3)])
def test_particles_follow_the_oracle(O, P, n, prec, nsteps):
    print("[gpu case] particles", n, prec, nsteps, flush=True)
    import oracle as orc
    g = O.Grid(n)
    u0 = g.field(seed=21)
    pp0 = orc.particles_uniform(n, 4096, 22)
    ro = O.Run(g, u0, 0.01, cfl=np.pi, kshift=1, kf_shell=3, shift_seed=21, particles=pp0)
    s = P.Solver(n, prec, viscos=0.01)
    s.put_state(u0)
    rg = P.Run(s, 0.01, np.pi, kf_shell=3, ekinf=ro.ekinf, particles=pp0)
    for _ in range(nsteps):
        gsh = ro.next_shifts()
        ro.step(gsh)
        rg.step(gsh)
    got = s.lag_get()
    tol = 1e-11 if prec == 8 else 1e-6
    assert rel(got[3:6], ro.pp[3:6]) < tol  # velocities at the predicted positions
    # positions in grid units: the displacement is what carries information
    disp_g, disp_o = got[0:3] - pp0[0:3], ro.pp[0:3] - pp0[0:3]
    assert np.abs(disp_o).max() > 0.05
    assert rel(disp_g, disp_o) < tol
    assert rel(got[6:9], ro.pp[6:9]) < 1e-12 if prec == 8 else True
    s.close()


def test_particles_512_fp32_2e20(O, P):
    """512^3 FP32, 2^20 particles uniform in the box: the spline velocity of
    the seeded field on a shifted grid at every particle, vs the oracle"""
    print("[gpu case] particles 512^3 fp32 2^20", flush=True)
    import oracle as orc
    n, npart = 512, 1 << 20
    g = O.Grid(n)
    u0 = g.field(seed=122929298)
    gsh = np.array([0.31, 0.72, 0.05, 0.81, 0.22, 0.55])
    pp = orc.particles_uniform(n, npart, 123)
    s = P.Solver(n, 4, viscos=0.01)
    s.put_state(u0)
    s.set_shifts(gsh)
    s.lag_put(pp[:3])
    s.lag_spline()
    s.lag_interp()
    got = s.lag_get()
    s.close()
    c = orc.velocity_spline(g, u0, gsh)
    orc.lag_interp(c, gsh[:3], pp)
    err = rel(got[3:6], pp[3:6])
    print("rel L2 of the particle velocities:", err, flush=True)
    assert err < 1e-6
    assert np.array_equal(got[0:3], pp[0:3])


def test_particle_call_order(O, P):
    print("[gpu case] particle errors", flush=True)
    import oracle as orc
    s = P.Solver(32, 8)
    try:
        s.put_state(O.Grid(32).field(seed=1))
        with pytest.raises(P.PsdnsError, match="no particles"):
            s.lag_interp()
        s.lag_put(orc.particles_uniform(32, 10, 1)[:3])
        with pytest.raises(P.PsdnsError, match="lag_spline first"):
            s.lag_interp()
        s.lag_spline()
        s.kxtran(1)  # a work-array call discards the coefficients
        with pytest.raises(P.PsdnsError, match="lag_spline first"):
            s.lag_interp()
        s.lag_spline()
        s.lag_interp()
        with pytest.raises(P.PsdnsError, match="pstep"):
            s.lag_advpos(3, 0.01, 0.5, 0.5)
    finally:
        s.close()
