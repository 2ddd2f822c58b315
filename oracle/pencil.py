"""Checkpoint files in the reference's pencil format, restated in numpy.

TEST INFRASTRUCTURE ONLY: tests/ use it as the checker of psdns_outpen /
psdns_inpen.  The GPU library never uses it.

Format (one task, iproc = jproc = 1; PSDNS/homogeneous/outpen_write.f:59-76,
read back by PSDNS/homogeneous/inpen_read.f with slabpenc = 2): Fortran
sequential-unformatted records, big-endian (-fconvert=big-endian,
PSDNS/makefile:251), each a 4-byte byte count, the payload, the count again.
Per variable i = 1 .. nv:
    (i, nx, ny, nz)                              int32 x 4
    (xist, xisz, zjst, zjsz, 1, ny)              int32 x 6  = (1, n/2, 1, n, 1, n)
    for z = 1 .. n, z != n/2+1 (outpen_write.f:68-75):
        buf(1:n/2, 1:n/2, z)      x fastest
        buf(1:n/2, n/2+2:n, z)
buf(x, y, z) is the cylinder value uny(y, a) of column a = (x, z), zero off
the cylinder (PSDNS/kxcomm1.f:586-612).

The pinning of this restatement: tests compare its bytes with a Fortran
restatement (oracle/pencil_write.f90) whose records come from the Fortran
runtime's own unformatted I/O, and check single values at offsets computed
from the record arithmetic.
"""
import numpy as np


def pencil(uny_var, n, col_x, col_z):
    """Cylinder (num_al, n) -> buf as [z][y][x] (x < n/2), zero off the cylinder."""
    buf = np.zeros((n, n, n // 2), np.complex128)
    buf[col_z, :, col_x] = uny_var
    return buf


def _rows(n):
    return np.arange(n // 2), np.arange(n // 2 + 1, n)


def _record(payload):
    m = np.array([len(payload)], ">i4").tobytes()
    return m + payload + m


def outpen_bytes(uny, n, col_x, col_z, file_prec=8):
    """File contents for uny of shape (nv, num_al, n)."""
    ct = ">c16" if file_prec == 8 else ">c8"
    lo, hi = _rows(n)
    out = []
    for i, var in enumerate(np.asarray(uny)):
        out.append(_record(np.array([i + 1, n, n, n], ">i4").tobytes()))
        out.append(_record(np.array([1, n // 2, 1, n, 1, n], ">i4").tobytes()))
        buf = pencil(var.astype(np.complex128), n, col_x, col_z)
        for z in range(n):
            if z == n // 2:
                continue
            out.append(_record(buf[z, lo, :].astype(ct).tobytes()))
            out.append(_record(buf[z, hi, :].astype(ct).tobytes()))
    return b"".join(out)


def write_outpen(path, uny, n, col_x, col_z, file_prec=8):
    with open(path, "wb") as f:
        f.write(outpen_bytes(uny, n, col_x, col_z, file_prec))


def _records(data):
    pos = 0
    while pos < len(data):
        m = int(np.frombuffer(data, ">i4", 1, pos)[0])
        payload = data[pos + 4:pos + 4 + m]
        m2 = int(np.frombuffer(data, ">i4", 1, pos + 4 + m)[0])
        assert m2 == m, "record markers differ"
        yield payload
        pos += 8 + m


def read_inpen(path, n, col_x, col_z, nv):
    """Cylinder (nv, num_al, n) complex128 and the file precision (4 or 8)."""
    data = open(path, "rb").read()
    recs = _records(data)
    lo, hi = _rows(n)
    uny = np.zeros((nv, len(col_x), n), np.complex128)
    fprec = None
    for i in range(nv):
        hdr = np.frombuffer(next(recs), ">i4")
        assert list(hdr) == [i + 1, n, n, n], hdr
        geo = np.frombuffer(next(recs), ">i4")
        assert list(geo) == [1, n // 2, 1, n, 1, n], geo
        buf = np.zeros((n, n, n // 2), np.complex128)
        for z in range(n):
            if z == n // 2:
                continue
            a = next(recs)
            if fprec is None:
                fprec = 8 if len(a) == 16 * (n // 2) * (n // 2) else 4
            ct = ">c16" if fprec == 8 else ">c8"
            buf[z, lo, :] = np.frombuffer(a, ct).reshape(len(lo), n // 2)
            buf[z, hi, :] = np.frombuffer(next(recs), ct).reshape(len(hi), n // 2)
        uny[i] = buf[col_z, :, col_x]
    return uny, fprec


def held(n, col_z):
    """Mask (num_al, n) of the cylinder values the file holds (not the ky = n/2
    row, not a kz = n/2 column)."""
    m = np.ones((len(col_z), n), bool)
    m[:, n // 2] = False
    m[np.asarray(col_z) == n // 2, :] = False
    return m
