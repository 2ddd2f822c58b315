"""EPFOR forcing increments through the C ABI against the FP64 oracle
(proc4a.f:588-640 with the efkz tally, proc4b.f:118-163, advanc_1..3), the
Fortran driver's epfor option (force.f construction + forcdt.f rate), and
the rule that a call which overwrites the work arrays between
stage_nonlinear and its update makes the update fail with PSDNS_ESTATE."""
import subprocess
import sys

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

TOL = {8: 1e-11, 4: 5e-6}
KFOR, KFORCE = 4, 3.2


def forcing_seq(O, nsteps, seed=31, amp=0.02):
    rng = np.random.default_rng(seed)
    return [O.epfor_increments(KFOR, KFORCE, rng, amp=amp) for _ in range(nsteps)]


@pytest.mark.parametrize("n,prec,nsteps", [(32, 8, 3), (64, 8, 3), (64, 4, 3)])
def test_epfor_steps_match_oracle(O, P, n, prec, nsteps):
    print("[gpu case] epfor rk2", n, prec, nsteps, flush=True)
    g = O.Grid(n)
    u0 = g.field(seed=13)
    fs = forcing_seq(O, nsteps)
    ro = O.Run(g, u0, 0.01, cfl=np.pi, kshift=1, shift_seed=13, epfor=(KFOR, KFORCE, lambda i: fs[i]))
    s = P.Solver(n, prec, viscos=0.01)
    s.put_state(u0)
    rg = P.Run(s, 0.01, np.pi, epfor=(KFOR, KFORCE, lambda i: fs[i]))
    m = g.mask()[None]
    for _ in range(nsteps):
        gsh = ro.next_shifts()
        ro.step(gsh)
        rg.step(gsh)
        efkz = s.epfor_tally()
        scale = np.abs(ro.efkz).max()
        assert scale > 0
        assert np.abs(efkz - ro.efkz).max() <= (1e-10 if prec == 8 else 1e-4) * scale
    assert rel(s.get_state(), ro.uny, m) < TOL[prec]
    # the forcing moved the state: a run without it differs by far more
    ro0 = O.Run(g, u0, 0.01, cfl=np.pi, kshift=1, shift_seed=13)
    for _ in range(nsteps):
        ro0.step()
    assert rel(ro0.uny, ro.uny, m) > 1e3 * TOL[prec]
    s.close()


@pytest.mark.parametrize("n,prec,nsteps", [(32, 8, 2), (64, 4, 2)])
def test_epfor_rk4_match_oracle(O, P, n, prec, nsteps):
    print("[gpu case] epfor rk4", n, prec, nsteps, flush=True)
    g = O.Grid(n)
    u0 = g.field(seed=17)
    fs = forcing_seq(O, nsteps, seed=5)
    ro = O.Run(g, u0, 0.02, cfl=np.pi, kshift=1, shift_seed=3, epfor=(KFOR, KFORCE, lambda i: fs[i]))
    s = P.Solver(n, prec, viscos=0.02)
    s.put_state(u0)
    rg = P.Run4(s, 0.02, np.pi, epfor=(KFOR, KFORCE, lambda i: fs[i]))
    m = g.mask()[None]
    for _ in range(nsteps):
        gsh = ro.next_shifts()
        ro.step_rk4(gsh)
        rg.step(gsh)
        scale = np.abs(ro.efkz).max()
        assert np.abs(s.epfor_tally() - ro.efkz).max() <= (1e-10 if prec == 8 else 1e-4) * scale
    assert rel(s.get_state(), ro.uny, m) < TOL[prec]
    s.close()


def test_epfor_needs_set_epfor_and_shape(O, P):
    print("[gpu case] epfor errors", flush=True)
    s = P.Solver(32, 8, viscos=0.01)
    s.put_state(O.Grid(32).field(seed=1))
    s.stage_nonlinear(1)
    f = forcing_seq(O, 1)[0]
    with pytest.raises(P.PsdnsError, match="set_epfor"):
        s.stage_update(1, 0.01, *s.step_factors(0.01), forcing=f)
    with pytest.raises(P.PsdnsError, match="kfor"):
        s.set_epfor(17, 3.0)
    with pytest.raises(P.PsdnsError, match="EPFOR forcing is not set"):
        s.epfor_tally()
    s.close()


def test_update_after_a_work_array_call_is_refused(O, P):
    """vg_moments (and the other work-array calls) between stage_nonlinear
    and stage_update discard the nonlinear terms: the update is refused"""
    print("[gpu case] estate after vg_moments", flush=True)
    s = P.Solver(32, 8, viscos=0.01)
    s.put_state(O.Grid(32).field(seed=2))
    s.stage_nonlinear(1)
    s.vg_moments()
    with pytest.raises(P.PsdnsError, match="without a preceding"):
        s.stage_update(1, 0.01, *s.step_factors(0.01))
    s.stage_nonlinear(1)
    s.kxtran(1)
    with pytest.raises(P.PsdnsError, match="without a preceding"):
        s.stage_update(1, 0.01, *s.step_factors(0.01))
    # the regular order still works afterwards
    s.stage_nonlinear(1)
    s.stage_update(1, 0.01, *s.step_factors(0.01))
    s.close()


def test_fortran_driver_epfor(O, P, tmp_path):
    print("[gpu case] fortran driver epfor 32^3 fp64", flush=True)
    n, nsteps = 32, 3
    g = O.Grid(n)
    u0 = g.field(seed=5)
    ro = O.Run(g, u0, 0.01, cfl=np.pi, kshift=1, shift_seed=5)
    shifts = ro.shift_sequence(nsteps)
    np.ascontiguousarray(u0, np.complex128).tofile(tmp_path / "ic.bin")
    np.ascontiguousarray(shifts, np.float64).tofile(tmp_path / "shifts.bin")
    cmd = [str(P.DRIVER_PATH), f"n={n}", "prec=8", f"nsteps={nsteps}", "cfl=3.14159265358979",
           "viscos=0.01", "kshift=1", f"kforce={KFORCE}", "epsfor=0.5", f"ic={tmp_path / 'ic.bin'}",
           f"shifts={tmp_path / 'shifts.bin'}", f"out={tmp_path / 'out.bin'}", f"log={tmp_path / 'log'}",
           f"epfor_out={tmp_path / 'for.bin'}"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    k2fo = 2 * KFOR - 1
    fs = np.fromfile(tmp_path / "for.bin", np.complex128).reshape(nsteps, 3, k2fo, k2fo, KFOR)
    assert np.abs(fs).max() > 0
    ro2 = O.Run(g, u0, 0.01, cfl=np.pi, kshift=1, shift_seed=5, epfor=(KFOR, KFORCE, lambda i: fs[i]))
    rates = []
    for k in range(nsteps):
        ro2.step(shifts[k])
        rates.append(0.5 * ro2.efkz.sum() / ro2.dt)
    out = np.fromfile(tmp_path / "out.bin", np.complex128).reshape(3, g.num_al, n)
    assert rel(out, ro2.uny, g.mask()[None]) < 1e-11
    got = [float(l.split()[-1]) for l in r.stdout.splitlines() if "forcing input" in l]
    assert len(got) == nsteps
    assert np.allclose(got, rates, rtol=1e-9, atol=1e-14)
