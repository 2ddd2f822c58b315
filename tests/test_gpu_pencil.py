"""Pencil checkpoint files through the C ABI (psdns_outpen / psdns_inpen)
against the oracle's restatement of homogeneous/outpen_write.f:59-76 and
inpen_read.f: byte-identical files, bit-exact reads, the restart running
the same step as put_state, and the error paths."""
import sys

import numpy as np
import pytest

from conftest import random_cylinder

pytestmark = pytest.mark.gpu

sys.path.insert(0, str(__import__("pathlib").Path(__file__).resolve().parent.parent / "oracle"))
import pencil  # noqa: E402


def banner(*a):
    print("[gpu case]", *a, flush=True)


@pytest.mark.parametrize("prec,fprec,nc", [(8, 8, 0), (8, 4, 1), (4, 4, 2), (4, 8, 0)])
def test_outpen_bytes_match_oracle(P, O, tmp_path, prec, fprec, nc):
    banner("outpen", prec, fprec, nc)
    n = 32
    g = O.Grid(n)
    s = P.Solver(n, prec=prec, nc=nc)
    u = random_cylinder(g, 3 + nc, 40 + nc, hermitian=False)
    s.put_state(u)
    st = s.get_state().astype(np.complex128)
    s.outpen(tmp_path / "outpenc.p0", fprec)
    got = (tmp_path / "outpenc.p0").read_bytes()
    assert got == pencil.outpen_bytes(st, n, g.col_x, g.col_z, fprec)
    s.close()


@pytest.mark.parametrize("prec,fprec", [(8, 8), (8, 4), (4, 4), (4, 8)])
def test_inpen_reads_oracle_file(P, O, tmp_path, prec, fprec):
    banner("inpen", prec, fprec)
    n, nc = 32, 1
    g = O.Grid(n)
    rng = np.random.default_rng(5)
    u = rng.standard_normal((3 + nc, g.num_al, n)) + 1j * rng.standard_normal((3 + nc, g.num_al, n))
    pencil.write_outpen(tmp_path / "f.p0", u, n, g.col_x, g.col_z, fprec)
    want, _ = pencil.read_inpen(tmp_path / "f.p0", n, g.col_x, g.col_z, 3 + nc)
    s = P.Solver(n, prec=prec, nc=nc)
    s.put_state(np.ones((3 + nc, g.num_al, n)))  # overwritten everywhere, held or not
    s.inpen(tmp_path / "f.p0")
    got = s.get_state()
    assert np.array_equal(got, want.astype(s.cdtype))
    # u = un after the read, as after put_state
    assert np.array_equal(s.get_field(P.FIELD_U, 3), got[:3])
    s.close()


@pytest.mark.parametrize("n,prec", [(32, 8), (64, 8), (128, 8), (64, 4)])
def test_restart_steps_like_put_state(P, O, tmp_path, n, prec):
    """outpen -> inpen into a fresh context, then one RK2 step on each:
    identical to the bit."""
    banner("restart", n, prec)
    g = O.Grid(n)
    u0 = g.field(seed=9)
    gsh = [0.1, 0.2, 0.3, 0.6, 0.7, 0.8]
    a = P.Solver(n, prec=prec, viscos=0.01)
    a.put_state(u0)
    P.Run(a, 0.01, np.pi).step(gsh)
    a.outpen(tmp_path / "chk.p0", prec)
    b = P.Solver(n, prec=prec, viscos=0.01)
    b.inpen(tmp_path / "chk.p0")
    assert np.array_equal(b.get_state(), a.get_state())
    ra, rb = P.Run(a, 0.01, np.pi), P.Run(b, 0.01, np.pi)
    ra.step(gsh)
    rb.step(gsh)
    assert ra.dt == rb.dt
    assert np.array_equal(a.get_state(), b.get_state())
    a.close()
    b.close()


def test_pencil_errors(P, O, tmp_path):
    banner("pencil errors")
    n = 32
    g = O.Grid(n)
    s = P.Solver(n, prec=8)
    s.put_state(g.field(seed=1))
    with pytest.raises(P.PsdnsError, match="error -1"):
        s.outpen(tmp_path / "x.p0", 2)
    with pytest.raises(P.PsdnsError, match="error -5"):
        s.outpen(tmp_path / "no_such_dir" / "x.p0", 8)
    with pytest.raises(P.PsdnsError, match="error -5"):
        s.inpen(tmp_path / "missing.p0")
    # another grid's file
    g16 = O.Grid(16)
    pencil.write_outpen(tmp_path / "n16.p0", np.zeros((3, g16.num_al, 16)), 16, g16.col_x, g16.col_z, 8)
    with pytest.raises(P.PsdnsError, match="error -1"):
        s.inpen(tmp_path / "n16.p0")
    # too few variables, and a truncated last record
    s.outpen(tmp_path / "ok.p0", 8)
    data = (tmp_path / "ok.p0").read_bytes()
    (tmp_path / "short.p0").write_bytes(data[: len(data) * 2 // 3])
    with pytest.raises(P.PsdnsError, match="error -5"):
        s.inpen(tmp_path / "short.p0")
    (tmp_path / "cut.p0").write_bytes(data[:-10])
    with pytest.raises(P.PsdnsError, match="error -5"):
        s.inpen(tmp_path / "cut.p0")
    # the state is still usable after the failed reads
    s.inpen(tmp_path / "ok.p0")
    s.close()
