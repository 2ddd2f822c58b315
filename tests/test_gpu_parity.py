"""GPU parity: each device entry point vs the FP64 oracle on the same seeded
inputs (through the C ABI).  Tolerances: FP64 1e-12 relative L2 (a few
1e-13 for single routines); FP32 2e-6 (one routine) / stated per test.
"""
import sys

import numpy as np
import pytest

from conftest import kgrid, random_cylinder, rel

pytestmark = pytest.mark.gpu

TOL = {8: 1e-12, 4: 3e-6}
CASES = [(n, p) for n in (32, 64, 128) for p in (8, 4)]


def banner(*a):
    print("[gpu case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctxs(P):
    cache = {}

    def get(n, prec):
        if (n, prec) not in cache:
            cache[(n, prec)] = P.Solver(n, prec, viscos=0.01)
        return cache[(n, prec)]

    yield get
    for s in cache.values():
        s.close()


@pytest.mark.parametrize("n,prec", CASES)
def test_kxtran(O, P, ctxs, n, prec):
    banner("kxtran", n, prec)
    g, s = O.Grid(n), ctxs(n, prec)
    u = random_cylinder(g, 3, 10 + n)
    s.put_field(P.FIELD_U, u)
    s.kxtran(3)
    got = s.get_field(P.FIELD_UX, 3)
    assert rel(got, g.kxtran(u, 3)) < TOL[prec]


@pytest.mark.parametrize("n,prec", CASES)
def test_xktran(O, P, ctxs, n, prec):
    banner("xktran", n, prec)
    g, s = O.Grid(n), ctxs(n, prec)
    ux = np.random.default_rng(n).standard_normal((4, n, n, n))
    s.put_field(P.FIELD_UX, ux)
    s.xktran(4)
    got = s.get_field(P.FIELD_U, 4)
    assert rel(got, g.xktran(ux, 4)) < TOL[prec]


@pytest.mark.parametrize("n,prec", CASES)
def test_phase1(O, P, ctxs, n, prec):
    banner("phase1", n, prec)
    g, s = O.Grid(n), ctxs(n, prec)
    gsh = np.array([0.13, 0.77, 0.41, 0.63, 0.27, 0.91])
    g.set_shifts(gsh)
    s.set_shifts(gsh)
    for kstep in (1, 2):
        u = g.zeros_cyl(5)
        u[:3] = random_cylinder(g, 3, kstep, hermitian=False) + 1.0  # off-mask entries nonzero
        s.put_field(P.FIELD_U, u[:3])
        s.phase1(kstep)
        g.phshift_inplace(u, 3, kstep)
        assert rel(s.get_field(P.FIELD_U, 3), u[:3]) < TOL[prec] / 4


@pytest.mark.parametrize("n,prec", CASES)
def test_itransform_vel(O, P, ctxs, n, prec):
    banner("itransform_vel", n, prec)
    g, s = O.Grid(n), ctxs(n, prec)
    u = g.zeros_cyl(5)
    u[:3] = random_cylinder(g, 3, 3)
    s.put_field(P.FIELD_U, u[:3])
    s.itransform_vel()
    assert rel(s.get_field(P.FIELD_UX, 3), g.itransform_vel(u)[:3]) < TOL[prec]


@pytest.mark.parametrize("n,prec", CASES)
def test_realspace_vel(O, P, ctxs, n, prec):
    banner("realspace_vel", n, prec)
    g, s = O.Grid(n), ctxs(n, prec)
    ux = np.zeros((5, n, n, n))
    ux[:3] = np.random.default_rng(4).standard_normal((3, n, n, n))
    ux = ux.astype(s.rdtype).astype(np.float64)  # same inputs at either precision
    s.put_field(P.FIELD_UX, ux)
    vm = s.realspace_vel(1)
    vref = g.realspace_vel(ux, 1)
    assert rel(s.get_field(P.FIELD_UX, 5), ux) < TOL[prec] / 10
    assert vm == pytest.approx(vref, rel=1e-12 if prec == 8 else 1e-6)


@pytest.mark.parametrize("n,prec", CASES)
def test_transform(O, P, ctxs, n, prec):
    banner("transform", n, prec)
    g, s = O.Grid(n), ctxs(n, prec)
    ux = np.random.default_rng(5).standard_normal((5, n, n, n))
    s.put_field(P.FIELD_UX, ux)
    s.transform()
    ref = g.transform(ux)
    got = s.get_field(P.FIELD_U, 4)
    assert rel(got, ref[:4]) < TOL[prec]


@pytest.mark.parametrize("n,prec", CASES)
def test_wavespace_vel(O, P, ctxs, n, prec):
    banner("wavespace_vel", n, prec)
    g, s = O.Grid(n), ctxs(n, prec)
    gsh = np.array([0.2, 0.5, 0.8, 0.7, 0.0, 0.3])
    g.set_shifts(gsh)
    s.set_shifts(gsh)
    for kstep in (1, 2):
        u = random_cylinder(g, 5, 20 + kstep, hermitian=False) + 0.5
        s.put_field(P.FIELD_U, u)
        s.wavespace_vel(kstep, 0.0123)
        g.wavespace_vel(u, kstep, 0.0123)
        assert rel(s.get_field(P.FIELD_U, 5), u) < TOL[prec] / 4


@pytest.mark.parametrize("n,prec", CASES)
def test_proc4a_proc4b(O, P, ctxs, n, prec):
    banner("proc4a/proc4b", n, prec)
    g, s = O.Grid(n), ctxs(n, prec)
    xf, yf, zf = g.step_factors(0.02, 0.031)
    u = g.zeros_cyl(5)
    u[:3] = random_cylinder(g, 3, 30)
    un = random_cylinder(g, 3, 31)
    s.put_field(P.FIELD_U, u[:3])
    s.put_field(P.FIELD_UN, un)
    s.proc4a(xf, yf, zf)
    g.proc4a(u, un, xf, yf, zf)
    m = g.mask()[None]
    assert rel(s.get_field(P.FIELD_U, 3), u[:3], m) < TOL[prec] / 4
    assert rel(s.get_field(P.FIELD_UN, 3), un, m) < TOL[prec] / 4
    u[:3] = random_cylinder(g, 3, 32)
    s.put_field(P.FIELD_U, u[:3])
    s.put_field(P.FIELD_UN, un)
    s.proc4b()
    g.proc4b(u, un)
    assert rel(s.get_field(P.FIELD_U, 3), u[:3], m) < TOL[prec] / 4
    assert rel(s.get_field(P.FIELD_UN, 3), un, m) < TOL[prec] / 4


@pytest.mark.parametrize("n,prec", [(64, 8), (64, 4), (128, 4)])
def test_init_field_and_stats(O, P, ctxs, n, prec):
    banner("init_field/shell_stats", n, prec)
    g, s = O.Grid(n), ctxs(n, prec)
    s.init_field(seed=122929298, kpeak=3.0, tke=0.5)
    ref = g.field(seed=122929298, kpeak=3.0, tke=0.5)
    got = s.get_state()
    assert rel(got, ref) < (1e-13 if prec == 8 else 1e-6)
    ek, dk, tke, eps = s.shell_stats(n)
    rek, rdk, rtke, reps = g.spectrum(ref, 0.01)
    assert tke == pytest.approx(0.5, rel=1e-12 if prec == 8 else 1e-6)
    assert eps == pytest.approx(reps, rel=1e-12 if prec == 8 else 1e-6)
    assert np.allclose(ek, rek, rtol=1e-11 if prec == 8 else 1e-5, atol=1e-14)
    assert np.allclose(dk, rdk, rtol=1e-11 if prec == 8 else 1e-5, atol=1e-14)


def oracle_stage(g, uy, uny, kstep, dt=None, viscos=0.01, cfl=np.pi):
    w = g.zeros_cyl(5)
    w[:3] = uy[:3]
    g.phshift_inplace(w, 3, kstep)
    ux = g.itransform_vel(w)
    vm = g.realspace_vel(ux, kstep)
    nl = g.transform(ux)
    if kstep == 1:
        dt = cfl / vm
    xf, yf, zf = g.step_factors(viscos, dt)
    g.wavespace_vel(nl, kstep, dt / 2)
    if kstep == 1:
        g.proc4a(nl, uny, xf, yf, zf)
    else:
        g.proc4b(nl, uny)
    return nl, uny, vm, dt


@pytest.mark.parametrize("peek", [True, False], ids=["peek", "straight"])
@pytest.mark.parametrize("n,prec", CASES)
def test_fused_stages(O, P, ctxs, n, prec, peek):
    """Both stages against the oracle.  With `peek` the predictor is read back
    between the stages, which runs stage 1's update on its own; without it
    stage_nonlinear(2) follows stage_update(1) directly and runs that update
    in one kernel with its own phase shift and inverse y pass."""
    banner(f"stage_nonlinear/stage_update peek={peek}", n, prec)
    g, s = O.Grid(n), ctxs(n, prec)
    gsh = np.array([0.35, 0.15, 0.65, 0.85, 0.65, 0.15])
    g.set_shifts(gsh)
    s.set_shifts(gsh)
    u0 = g.field(seed=77)
    s.put_state(u0)
    vm = s.stage_nonlinear(1)
    uy, uny, vref, dt = oracle_stage(g, u0.copy(), u0.copy(), 1)
    assert vm == pytest.approx(vref, rel=1e-12 if prec == 8 else 2e-6)
    xf, yf, zf = g.step_factors(0.01, dt)
    s.stage_update(1, dt, xf, yf, zf)
    m = g.mask()[None]
    if peek:
        assert rel(s.get_field(P.FIELD_U, 3), uy[:3], m) < TOL[prec]
        assert rel(s.get_state(), uny, m) < TOL[prec]
    s.stage_nonlinear(2)
    s.stage_update(2, dt, xf, yf, zf)
    uy2, uny2, _, _ = oracle_stage(g, uy, uny, 2, dt)
    assert rel(s.get_state(), uny2, m) < TOL[prec]
    assert rel(s.get_field(P.FIELD_U, 3), uy2[:3], m) < TOL[prec]


@pytest.mark.parametrize("n,prec", [(64, 8), (64, 4)])
def test_deferred_update_keeps_its_shifts(O, P, ctxs, n, prec):
    """stage_update(1) runs in the next call; new shifts set in between apply
    to stage 2 only (the deferred update still uses the stage-1 factors it
    was issued with)."""
    banner("deferred update + set_shifts", n, prec)
    g, s = O.Grid(n), ctxs(n, prec)
    gsh1 = np.array([0.35, 0.15, 0.65, 0.85, 0.65, 0.15])
    gsh2 = np.array([0.55, 0.25, 0.05, 0.45, 0.95, 0.75])
    g.set_shifts(gsh1)
    s.set_shifts(gsh1)
    u0 = g.field(seed=78)
    s.put_state(u0)
    s.stage_nonlinear(1)
    uy, uny, _, dt = oracle_stage(g, u0.copy(), u0.copy(), 1)
    xf, yf, zf = g.step_factors(0.01, dt)
    s.stage_update(1, dt, xf, yf, zf)
    g.set_shifts(gsh2)
    s.set_shifts(gsh2)
    s.stage_nonlinear(2)
    s.stage_update(2, dt, xf, yf, zf)
    uy2, uny2, _, _ = oracle_stage(g, uy, uny, 2, dt)
    m = g.mask()[None]
    assert rel(s.get_state(), uny2, m) < TOL[prec]
    assert rel(s.get_field(P.FIELD_U, 3), uy2[:3], m) < TOL[prec]


def test_stage_order_is_enforced(P, ctxs):
    s = ctxs(32, 8)
    with pytest.raises(P.PsdnsError, match="without a preceding"):
        s.stage_update(2, 0.01, *s.step_factors(0.01))
    with pytest.raises(P.PsdnsError, match="EPFOR"):
        s.L.psdns_stage_nonlinear(s.h, 1, None)
        xf, yf, zf = (np.ascontiguousarray(f) for f in s.step_factors(0.01))
        import ctypes
        dp = ctypes.POINTER(ctypes.c_double)
        rc = s.L.psdns_stage_update(s.h, 1, 0.01, xf.ctypes.data_as(dp), yf.ctypes.data_as(dp),
                                    zf.ctypes.data_as(dp), ctypes.c_void_p(1))
        s._chk(rc)


@pytest.mark.parametrize("n,prec", [(32, 8), (64, 8), (64, 4), (128, 4)])
def test_vg_moments(O, P, ctxs, n, prec):
    """Moments 1..8 of du/dx and dw/dz (vgmomt.f:68-143) vs the oracle on a
    seeded field; FP64 sums on both sides."""
    banner("vg_moments", n, prec)
    g, s = O.Grid(n), ctxs(n, prec)
    u0 = g.field(seed=91 + n)
    s.put_state(u0)
    du, dw = s.vg_moments()
    rdu, rdw = g.vg_moments(u0)
    tol = 1e-11 if prec == 8 else 2e-5
    for got, ref in ((du, rdu), (dw, rdw)):
        # relative, with the m-th power of the rms as the floor for the
        # near-zero odd moments
        scale = np.maximum(np.abs(ref), np.abs(ref[1]) ** (np.arange(1, 9) / 2.0))
        assert np.all(np.abs(got - ref) <= tol * scale), (got, ref)
    assert du[1] > 0 and dw[1] > 0
