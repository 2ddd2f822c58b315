"""The pencil checkpoint format of the oracle (oracle/pencil.py) pinned on the
CPU: byte-identical to the files the Fortran runtime writes from a Fortran
restatement of homogeneous/outpen_write.f:68-76 (oracle/pencil_write.f90,
compiled here with amdflang, big-endian records as -fconvert=big-endian
gives, makefile:251), single values found at offsets computed from the
record arithmetic, and the reader's round trip."""
import pathlib
import shutil
import subprocess
import sys

import numpy as np
import pytest

REPO = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "oracle"))
import pencil  # noqa: E402

FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"


def random_state(grid, nv, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nv, grid.num_al, grid.n)) + 1j * rng.standard_normal((nv, grid.num_al, grid.n))


@pytest.fixture(scope="module")
def fortran_writer(tmp_path_factory):
    if not pathlib.Path(FC).exists():
        pytest.skip("amdflang not present")
    exe = tmp_path_factory.mktemp("fw") / "pencil_write"
    subprocess.run([FC, "-O1", str(REPO / "oracle" / "pencil_write.f90"), "-o", str(exe)], check=True,
                   capture_output=True, timeout=300)
    return exe


@pytest.mark.parametrize("n,nv,fprec", [(16, 3, 8), (16, 4, 4), (32, 5, 8)])
def test_numpy_writer_matches_fortran_records(O, fortran_writer, tmp_path, n, nv, fprec):
    g = O.Grid(n)
    u = random_state(g, nv, 11 + n + nv)
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(np.array([n, g.num_al, nv, fprec], np.int32).tobytes())
        f.write(g.col_x.astype(np.int32).tobytes())
        f.write(g.col_z.astype(np.int32).tobytes())
        # Fortran uny(n, num_al, nv) == C order (nv, num_al, n)
        f.write(np.ascontiguousarray(u, np.complex128).tobytes())
    subprocess.run([str(fortran_writer), str(tmp_path / "in.bin"), str(tmp_path / "ref.p0")], check=True,
                   timeout=120)
    ref = (tmp_path / "ref.p0").read_bytes()
    assert pencil.outpen_bytes(u, n, g.col_x, g.col_z, fprec) == ref


def test_values_at_computed_offsets(O):
    """File offsets from the record arithmetic alone: header records of
    4 + 16 + 4 and 4 + 24 + 4 bytes, then per stored z plane a record of
    (n/2)(n/2) values and one of (n/2)(n/2 - 1), x fastest."""
    n, nv, cs = 16, 3, 16
    g = O.Grid(n)
    u = random_state(g, nv, 3)
    data = pencil.outpen_bytes(u, n, g.col_x, g.col_z, 8)
    nxh = n // 2
    rec_a, rec_b = nxh * (n // 2) * cs, nxh * (n // 2 - 1) * cs
    per_plane = 8 + rec_a + 8 + rec_b
    per_var = 24 + 32 + (n - 1) * per_plane
    assert len(data) == nv * per_var

    def offset(i, x, y, z):
        zp = z if z < n // 2 else z - 1  # z = n/2 is not stored
        o = i * per_var + 24 + 32 + zp * per_plane
        if y < n // 2:
            return o + 4 + (y * nxh + x) * cs
        return o + 8 + rec_a + 4 + ((y - n // 2 - 1) * nxh + x) * cs

    assert list(np.frombuffer(data, ">i4", 6, 0)) == [16, 1, n, n, n, 16]
    assert list(np.frombuffer(data, ">i4", 8, 24)) == [24, 1, n // 2, 1, n, 1, n, 24]
    rng = np.random.default_rng(0)
    for _ in range(200):
        i, a, y = rng.integers(nv), rng.integers(g.num_al), rng.integers(n)
        x, z = int(g.col_x[a]), int(g.col_z[a])
        if y == n // 2 or z == n // 2:
            continue
        v = np.frombuffer(data, ">c16", 1, offset(i, x, y, z))[0]
        assert v == u[i, a, y]
    # a column outside the cylinder is stored as zeros
    cols = set(zip(g.col_x.tolist(), g.col_z.tolist()))
    x, z = next((x, z) for x in range(nxh) for z in range(n) if (x, z) not in cols and z != n // 2)
    assert np.frombuffer(data, ">c16", 1, offset(1, x, 3, z))[0] == 0


@pytest.mark.parametrize("fprec", [8, 4])
def test_reader_round_trip(O, tmp_path, fprec):
    n, nv = 32, 4
    g = O.Grid(n)
    u = random_state(g, nv, 7)
    pencil.write_outpen(tmp_path / "f.p0", u, n, g.col_x, g.col_z, fprec)
    back, fp = pencil.read_inpen(tmp_path / "f.p0", n, g.col_x, g.col_z, nv)
    assert fp == fprec
    held = pencil.held(n, g.col_z)
    want = u if fprec == 8 else u.astype(np.complex64).astype(np.complex128)
    assert np.array_equal(back[:, held], want[:, held])
    assert not back[:, ~held].any()
    # the modes the file does not hold are off the reference mask anyway
    assert not (g.mask() & ~held).any()
