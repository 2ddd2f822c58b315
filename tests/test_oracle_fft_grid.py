"""Oracle pins (CPU): the 1-D transform contract and the cylinder geometry.

- transforms vs numpy's pocketfft (committed fixtures tests/golden/fft_vectors.npz
  and live calls), FFTW signs of epfftw.f:77,135,146, unnormalised;
- cylinder counts vs tests/golden/cylinder_counts.json (mpisetup.f:311-354,
  masks.f:67-77 restated);
- column order vs the next_xz state machine of mpisetup.f:710-732 iterated
  literally, and the mask vs the floating-point test of masks.f:67-77.
"""
import json
import pathlib

import numpy as np
import pytest

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


def test_fft_vs_committed_numpy_vectors(O):
    g = np.load(GOLD / "fft_vectors.npz")
    for n in (8, 16, 32, 64):
        x = g[f"c2c_in_{n}"]
        assert np.allclose(O.fft_c2c(x, -1), g[f"c2c_fwd_{n}"], rtol=0, atol=1e-12 * n)
        assert np.allclose(O.fft_c2c(x, +1), g[f"c2c_bwd_{n}"], rtol=0, atol=1e-12 * n)
        assert np.allclose(O.fft_r2c(g[f"r2c_in_{n}"]), g[f"r2c_out_{n}"], rtol=0, atol=1e-12 * n)
        assert np.allclose(O.fft_c2r(g[f"c2r_in_{n}"], n), g[f"c2r_out_{n}"], rtol=0, atol=1e-12 * n)


@pytest.mark.parametrize("n", [2, 4, 128, 1024])
def test_fft_vs_numpy_live(O, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((5, n)) + 1j * rng.standard_normal((5, n))
    assert np.abs(O.fft_c2c(x, -1) - np.fft.fft(x)).max() < 1e-12 * n
    assert np.abs(O.fft_c2c(x, +1) - np.fft.ifft(x) * n).max() < 1e-12 * n
    r = rng.standard_normal((5, n))
    assert np.abs(O.fft_r2c(r) - np.fft.rfft(r)).max() < 1e-12 * n


def test_c2r_discards_imaginary_dc_and_nyquist(O):
    n = 16
    rng = np.random.default_rng(3)
    X = rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)
    ref = np.fft.irfft(X, n) * n  # numpy drops Im X[0] and Im X[n/2]
    assert np.abs(O.fft_c2r(X[None], n)[0] - ref).max() < 1e-12
    Xr = X.copy()
    Xr[0] = Xr[0].real
    Xr[-1] = Xr[-1].real
    assert np.abs(O.fft_c2r(Xr[None], n)[0] - ref).max() < 1e-12


@pytest.mark.parametrize("n", [64, 128, 256, 512, 1024])
def test_cylinder_counts(O, n):
    want = json.loads((GOLD / "cylinder_counts.json").read_text())["counts"][str(n)]
    g = O.Grid(n)
    assert g.max_al_x == want["max_al_x"]
    assert g.num_al == want["num_al"]
    if "mask_modes" in want:
        assert g.mask_count() == want["mask_modes"]
        assert int(g.mask().sum()) == want["mask_modes"]


def _next_xz_columns(n, max_al_x, cut_z):
    """mpisetup.f:710-732 iterated literally (1-based x, z)"""
    nzhp = n // 2 + 1
    cols = []
    x, z = 1, 1
    total = sum(n - 1 - 2 * c for c in cut_z)
    for _ in range(total):
        cols.append((x - 1, z - 1))
        c = cut_z[x - 1]
        z1, z2 = nzhp - c - 1, nzhp + c + 1
        if z < z1:
            z += 1
        elif z < z2:
            z = z2
        elif z < n:
            z += 1
        else:
            z, x = 1, x + 1
    return cols


@pytest.mark.parametrize("n", [32, 64, 128])
def test_column_order_matches_next_xz(O, n):
    g = O.Grid(n)
    # cut_z from the mpisetup.f:331-332 formula evaluated independently
    cut = []
    for x in range(g.max_al_x):
        rc = n * (1.0 - 2.0 * np.sqrt(2.0 / 9.0 - (x / n) ** 2))
        cut.append(int(rc / 2))
    assert list(g.cut_z) == cut
    cols = _next_xz_columns(n, g.max_al_x, cut)
    assert len(cols) == g.num_al
    assert list(zip(g.col_x.tolist(), g.col_z.tolist())) == cols


@pytest.mark.parametrize("n", [32, 64])
def test_mask_matches_floating_point_test(O, n):
    g = O.Grid(n)
    m = g.mask()
    rnxi = 1.0 / n
    two9 = 2.0 / 9.0
    for a in range(0, g.num_al, 7):
        x, z = int(g.col_x[a]), int(g.col_z[a])
        kz = z if z <= n // 2 else z - n
        for y in range(n):
            ky = y if y <= n // 2 else y - n
            sqk = (ky * rnxi) ** 2 + ((kz * rnxi) ** 2 + (x * rnxi) ** 2)
            want = not (sqk > two9 or y == n // 2 or z == n // 2)
            assert m[a, y] == want
            assert bool(O.lib().orc_mask(g.h, y, a)) == want


def test_every_mask_mode_lives_in_the_cylinder(O):
    # every (kx, ky, kz) in the sphere with kx < n/2 has its (x, z) column in the cylinder
    n = 64
    g = O.Grid(n)
    cols = set(zip(g.col_x.tolist(), g.col_z.tolist()))
    for x in range(n // 2):
        for z in range(n):
            kz = z if z <= n // 2 else z - n
            if 9 * (x * x + kz * kz) <= 2 * n * n and z != n // 2:
                assert (x, z) in cols
