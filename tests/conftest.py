"""Shared fixtures: the package (loaded as `psdns_amd`), the oracle, helpers.

GPU tests are marked `gpu`; everything else runs on the CPU.  The oracle
(oracle/) is the checker only.
"""
import importlib.util
import pathlib
import sys

import numpy as np
import pytest

REPO = pathlib.Path(__file__).resolve().parent.parent
PKG_DIR = REPO / "nasa-bluewater-psdns-benchmark-minsky_amd"


def load_package():
    if "psdns_amd" in sys.modules:
        return sys.modules["psdns_amd"]
    spec = importlib.util.spec_from_file_location("psdns_amd", PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["psdns_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    if str(REPO / "oracle") not in sys.path:
        sys.path.insert(0, str(REPO / "oracle"))
    import oracle  # noqa: E402

    return oracle


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


@pytest.fixture(scope="session")
def P():
    return load_package()


@pytest.fixture(scope="session")
def O():
    return load_oracle()


def kgrid(n, col_x, col_z):
    """wavenumber arrays broadcastable over (num_al, n) cylinder entries"""
    kx = col_x.astype(np.float64)[:, None]
    kz = np.where(col_z > n // 2, col_z - n, col_z).astype(np.float64)[:, None]
    y = np.arange(n)
    ky = np.where(y > n // 2, y - n, y).astype(np.float64)[None, :]
    return kx, ky, kz


def rel(a, b, mask=None):
    a = np.asarray(a)
    b = np.asarray(b)
    if mask is not None:
        a = a * mask
        b = b * mask
    den = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0)


def random_cylinder(grid, nv, seed, hermitian=True):
    """seeded random cylinder data on the mask (kx=0 plane Hermitian so the
    field is the transform of a real field)"""
    rng = np.random.default_rng(seed)
    n = grid.n
    u = (rng.standard_normal((nv, grid.num_al, n)) + 1j * rng.standard_normal((nv, grid.num_al, n)))
    u *= grid.mask()[None]
    if hermitian:
        colz0 = {int(z): a for a, (x, z) in enumerate(zip(grid.col_x, grid.col_z)) if x == 0}
        for v in range(nv):
            a0 = colz0[0]
            u[v, a0, 0] = u[v, a0, 0].real
            for y in range(n // 2 + 1, n):
                u[v, a0, y] = np.conj(u[v, a0, n - y])
            for z in range(n // 2 + 1, n):
                if z in colz0 and (n - z) in colz0:
                    for y in range(n):
                        u[v, colz0[z], y] = np.conj(u[v, colz0[n - z], (n - y) % n])
    return np.ascontiguousarray(u)
