"""Pins of the SGS-closure oracle (oracle/sgs.py) by the mathematics: the
box filter's transfer function on single Fourier modes and a brute-force
triple sum, the subgrid stress of a single mode in closed form, spectral
gradients of a single mode, invariants of pure strain and pure rotation,
rotation invariance, the backward pass against central finite differences,
the first Adam step, and splitmix64's published first outputs."""
import numpy as np
import pytest

from conftest import load_oracle

load_oracle()  # puts oracle/ on sys.path
import sgs as SG  # noqa: E402


def h1(k, r, n):
    """transfer function of the 1-D (r + 1)-tap top-hat at wavenumber k on n points"""
    offs, w = SG.box_weights(r)
    return float(np.sum(w * np.cos(2 * np.pi * k * offs / n)))


def test_box_weights_sum_to_one():
    for r in (2, 4, 8):
        offs, w = SG.box_weights(r)
        assert abs(w.sum() - 1.0) < 1e-15 and len(w) == r + 1 and offs[0] == -r // 2


def test_box_filter_single_mode_closed_form():
    n, r = 32, 4
    x = 2 * np.pi * np.arange(n) / n
    Y, Z, X = np.meshgrid(x, x, x, indexing="ij")
    kx, ky, kz, a = 3, 2, 5, 0.7
    u = np.stack([a * np.cos(kx * X + ky * Y), np.zeros_like(X), np.sin(kz * Z)])
    ubar, tau = SG.box_filter(u, r)
    xc = X[::r, ::r, ::r], Y[::r, ::r, ::r], Z[::r, ::r, ::r]
    Xc, Yc, Zc = xc
    np.testing.assert_allclose(ubar[0], a * h1(kx, r, n) * h1(ky, r, n) * np.cos(kx * Xc + ky * Yc), atol=1e-13)
    np.testing.assert_allclose(ubar[2], h1(kz, r, n) * np.sin(kz * Zc), atol=1e-13)
    # tau_33 = filter(w^2) - wbar^2 with w^2 = (1 - cos 2 kz z) / 2
    t33 = 0.5 * (1 - h1(2 * kz, r, n) * np.cos(2 * kz * Zc)) - (h1(kz, r, n) * np.sin(kz * Zc)) ** 2
    np.testing.assert_allclose(tau[5], t33, atol=1e-13)
    # tau_13 = filter(u w) - ubar wbar: the product of modes in different directions
    c = np.cos(kx * X + ky * Y) * np.sin(kz * Z)
    f13 = a * h1(kx, r, n) * h1(ky, r, n) * h1(kz, r, n) * c[::r, ::r, ::r]
    np.testing.assert_allclose(tau[2], f13 - ubar[0] * ubar[2], atol=1e-13)
    assert np.abs(tau[2]).max() < 1e-13  # separable: the filter of the product is the product of the filters


def test_box_filter_brute_force_tiny():
    rng = np.random.default_rng(3)
    n, r = 8, 2
    u = rng.standard_normal((3, n, n, n))
    ubar, tau = SG.box_filter(u, r)
    offs, w = SG.box_weights(r)
    for (cy, cz, cx) in [(0, 0, 0), (1, 3, 2), (3, 3, 3)]:
        s = np.zeros(3)
        s2 = np.zeros((3, 3))
        for dy, wy in zip(offs, w):
            for dz, wz in zip(offs, w):
                for dx, wx in zip(offs, w):
                    v = u[:, (cy * r + dy) % n, (cz * r + dz) % n, (cx * r + dx) % n]
                    s += wy * wz * wx * v
                    s2 += wy * wz * wx * np.outer(v, v)
        np.testing.assert_allclose(ubar[:, cy, cz, cx], s, atol=1e-13)
        want = [s2[i, j] - s[i] * s[j] for i, j in SG.SYM]
        np.testing.assert_allclose(tau[:, cy, cz, cx], want, atol=1e-13)


def test_constant_field_has_no_stress():
    u = np.ones((3, 16, 16, 16)) * np.array([0.3, -1.0, 2.0])[:, None, None, None]
    ubar, tau = SG.box_filter(u, 4)
    assert np.abs(tau).max() < 1e-14
    np.testing.assert_allclose(ubar[:, 1, 2, 3], [0.3, -1.0, 2.0], atol=1e-14)


def test_gradient_single_mode(O):
    n = 32
    g = O.Grid(n)
    x = 2 * np.pi * np.arange(n) / n
    Y, Z, X = np.meshgrid(x, x, x, indexing="ij")
    # a solenoidal field of three modes: u = sin(2 z), v = cos(3 x + z), w = 0
    u = np.stack([np.sin(2 * Z), np.cos(3 * X + Z), np.zeros_like(X)])
    uy = g.xktran(u, 3)
    A = SG.gradient(g, uy)
    np.testing.assert_allclose(A[0, 2], 2 * np.cos(2 * Z), atol=1e-12)
    np.testing.assert_allclose(A[1, 0], -3 * np.sin(3 * X + Z), atol=1e-12)
    np.testing.assert_allclose(A[1, 2], -np.sin(3 * X + Z), atol=1e-12)
    for (i, j) in [(0, 0), (0, 1), (1, 1), (2, 0), (2, 1), (2, 2)]:
        assert np.abs(A[i, j]).max() < 1e-12


def test_invariants_pure_strain_and_rotation():
    a, b = 0.7, -0.2
    A = np.diag([a, b, -a - b])[:, :, None]
    lam, T = SG.invariants(A, 0.5)
    s2 = a * a + b * b + (a + b) ** 2
    s3 = a ** 3 + b ** 3 - (a + b) ** 3
    np.testing.assert_allclose(lam[:, 0], [1.0, 0.0, s3 / s2 ** 1.5, 0.0, 0.0], atol=1e-15)
    d2 = 0.25
    np.testing.assert_allclose(T[0, :, 0], d2 * np.sqrt(2 * s2) * np.array([a, 0, 0, b, 0, -a - b]), atol=1e-15)
    assert np.abs(T[1]).max() < 1e-15 and np.abs(T[3]).max() < 1e-15
    np.testing.assert_allclose(T[2, :, 0], d2 * (np.array([a * a, 0, 0, b * b, 0, (a + b) ** 2]) - s2 / 3 * np.array([1, 0, 0, 1, 0, 1])), atol=1e-15)
    # pure rotation about z: W = [[0, -w, 0], [w, 0, 0], [0, 0, 0]]
    w = 1.3
    A = np.array([[0, -w, 0], [w, 0, 0], [0, 0, 0.0]])[:, :, None]
    lam, T = SG.invariants(A, 1.0)
    np.testing.assert_allclose(lam[:, 0], [0.0, -1.0, 0.0, 0.0, 0.0], atol=1e-15)
    assert np.abs(T[:3]).max() < 1e-15
    # W^2 = diag(-w^2, -w^2, 0), tr = -2 w^2
    np.testing.assert_allclose(T[3, :, 0], [-w * w + 2 * w * w / 3, 0, 0, -w * w + 2 * w * w / 3, 0, 2 * w * w / 3], atol=1e-14)


def test_invariants_rotation_invariance():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((3, 3, 7))
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    Ar = np.einsum("ik,klp,jl->ijp", Q, A, Q)
    lam, T = SG.invariants(A, 0.3)
    lr, Tr = SG.invariants(Ar, 0.3)
    np.testing.assert_allclose(lr, lam, rtol=1e-12, atol=1e-13)

    def full(t):
        M = np.zeros((3, 3) + t.shape[1:])
        for c, (i, j) in enumerate(SG.SYM):
            M[i, j] = M[j, i] = t[c]
        return M

    for k in range(4):
        np.testing.assert_allclose(full(Tr[k]), np.einsum("ik,klp,jl->ijp", Q, full(T[k]), Q), atol=1e-12)
        M = full(T[k])
        assert np.abs(M[0, 0] + M[1, 1] + M[2, 2]).max() < 1e-13  # traceless


def _tiny_batch(seed, P=7):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((P, SG.NIN0)), rng.standard_normal((P, SG.NLAM)),
            rng.standard_normal((P, SG.NBASIS, 6)), rng.standard_normal((P, 6)))


def test_backward_matches_finite_differences():
    theta = SG.init_params(11)
    theta = theta + 0.05 * np.random.default_rng(12).standard_normal(theta.size)  # nonzero biases too
    X0, lam, T, tau = _tiny_batch(13)
    L, g = SG.loss_and_grad(theta, X0, lam, T, tau, 0.7)
    rng = np.random.default_rng(14)
    offs, o = [], 0
    for name, shape in SG.PARAM_SHAPES:
        cnt = int(np.prod(shape))
        offs += list(o + rng.choice(cnt, size=min(cnt, 6), replace=False))
        o += cnt
    h = 1e-6
    for i in offs:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (SG.loss_and_grad(tp, X0, lam, T, tau, 0.7)[0] - SG.loss_and_grad(tm, X0, lam, T, tau, 0.7)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(fd)), (i, fd, g[i])


def test_planted_backward_mistake_is_caught():
    """a dropped tanh derivative in one layer gives gradients the finite
    differences reject"""
    theta = SG.init_params(21)
    X0, lam, T, tau = _tiny_batch(22)
    p = SG.unpack(theta)
    L, g = SG.loss_and_grad(theta, X0, lam, T, tau, 1.0)
    # the W2 gradient without (1 - h2^2)
    _, (X0c, x1, h0, h1, h2, h3, gco) = SG.forward(theta, X0, lam, T)
    tau_hat = np.einsum("pn,pnc->pc", gco, T)
    dtau = 2.0 * SG.CW * (tau_hat - tau) / (X0.shape[0] * 1.0)
    dg = np.einsum("pc,pnc->pn", dtau, T)
    d3 = (dg @ p["W4"]) * (1 - h3 * h3)
    d2_wrong = d3 @ p["W3"]
    gw2_wrong = d2_wrong.T @ h1
    o = SG.PARAM_SHAPES
    off = sum(int(np.prod(s)) for name, s in o[:4])
    gw2 = g[off:off + 64 * 64].reshape(64, 64)
    assert np.abs(gw2_wrong - gw2).max() > 1e-3 * np.abs(gw2).max()


def test_first_adam_step_is_a_signed_step():
    rng = np.random.default_rng(8)
    th = rng.standard_normal(50)
    g = rng.standard_normal(50)
    t1, m, v = SG.adam(th, np.zeros(50), np.zeros(50), g, 1, 1e-3)
    # m_hat = g, v_hat = g^2: the step is lr g / (|g| + eps)
    np.testing.assert_allclose(t1, th - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_splitmix64_known_outputs():
    # SplitMix64 (Steele, Lea & Flood 2014) from state 0: the first outputs
    assert SG.splitmix64(0) == 0xE220A8397B1DCDAF
    assert SG.splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4
    idx = SG.sample_points(7, 3, 100, 2, 32 ** 3)
    assert idx.min() >= 0 and idx.max() < 2 * 32 ** 3 and len(set(idx.tolist())) > 90
    assert np.array_equal(idx, SG.sample_points(7, 3, 100, 2, 32 ** 3))


def test_param_count():
    assert SG.NPARAM == 12164
