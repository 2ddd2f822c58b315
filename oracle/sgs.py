"""CPU FP64 oracle of the learned subgrid-stress (SGS) closure.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline leg
as the checker; the GPU library never uses it.

The reference has no counterpart (PSDNS is a DNS code without an LES
closure); the algorithm restated here is the GPU library's own design, so
each function is pinned by the mathematics instead of a reference file:
  * box_filter      - top-hat of width r fine cells, trapezoidal end weights
                      (separable (r + 1)-tap kernel [1/2, 1, .., 1, 1/2] / r)
                      centred on every r-th fine point; tau_ij = filter(u_i
                      u_j) - filter(u_i) filter(u_j).  Pinned by exactness on
                      constants and linear fields, and by the brute-force
                      triple sum on a tiny grid.
  * gradient        - A_ij = du_i/dx_j of the spectral (cylinder) state by
                      i k_j, through the oracle's kxtran; pinned by a single
                      Fourier mode (closed form).
  * invariants      - the five invariants of S, W (tr S^2, tr W^2, tr S^3,
                      tr W^2 S, tr W^2 S^2) over powers of q = |S|^2 + |W|^2
                      and the tensor basis delta^2 {|S| S, S W - W S,
                      S^2 - tr(S^2)/3 I, W^2 - tr(W^2)/3 I}; pinned by pure
                      strain / pure rotation closed forms and by rotation
                      invariance.
  * network         - conv 3x3x3 (5 -> 16, tanh) on the invariant field,
                      MLP 21 -> 64 -> 64 -> 64 (tanh) -> 4 basis coefficients;
                      backward pinned by central finite differences in FP64.
  * adam            - Kingma & Ba (2015) Algorithm 1 with bias correction.
  * sample_points   - the splitmix64 finaliser (Steele, Lea & Flood 2014) of
                      a counter, the minibatch index map the GPU uses.

Arrays: physical fields (nv, n, n, n) [v][y][z][x] as the oracle's ux;
symmetric tensors as 6 components (11, 12, 13, 22, 23, 33).
"""
import numpy as np

SYM = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))
CW = np.array([1.0, 2.0, 2.0, 1.0, 2.0, 1.0])  # Frobenius weights of the 6 components
NLAM, NCONV, NHID, NBASIS = 5, 16, 64, 4
NTAP = 27
NIN0 = NLAM * NTAP       # conv inputs per point (135): k = ch * 27 + d
NIN1 = NLAM + NCONV      # first MLP layer inputs (21)
QEPS = 1e-30             # floor of q = |S|^2 + |W|^2 (zero-gradient points)

PARAM_SHAPES = (("W0", (NCONV, NIN0)), ("b0", (NCONV,)), ("W1", (NHID, NIN1)), ("b1", (NHID,)),
                ("W2", (NHID, NHID)), ("b2", (NHID,)), ("W3", (NHID, NHID)), ("b3", (NHID,)),
                ("W4", (NBASIS, NHID)), ("b4", (NBASIS,)))
NPARAM = sum(int(np.prod(s)) for _, s in PARAM_SHAPES)  # 12164


# ---------------------------------------------------------------------------
# parameters
# ---------------------------------------------------------------------------
def unpack(theta):
    out, o = {}, 0
    for name, shape in PARAM_SHAPES:
        cnt = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + cnt]).reshape(shape)
        o += cnt
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], np.float64).ravel() for name, _ in PARAM_SHAPES])


def init_params(seed):
    """Glorot-uniform weights, zero biases; the output layer scaled by 0.1
    so that the untrained closure starts near zero stress"""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in PARAM_SHAPES:
        if name[0] == "W":
            a = np.sqrt(6.0 / (shape[0] + shape[1]))
            p[name] = rng.uniform(-a, a, shape) * (0.1 if name == "W4" else 1.0)
        else:
            p[name] = np.zeros(shape)
    return pack(p)


# ---------------------------------------------------------------------------
# filter, gradient, invariants
# ---------------------------------------------------------------------------
def box_weights(r):
    assert r >= 2 and r % 2 == 0
    w = np.full(r + 1, 1.0 / r)
    w[0] = w[-1] = 0.5 / r
    return np.arange(-r // 2, r // 2 + 1), w


def _filter_sample(f, r):
    offs, w = box_weights(r)
    g = f
    for ax in range(3):
        g = sum(wi * np.roll(g, -o, axis=ax) for o, wi in zip(offs, w))
    return g[::r, ::r, ::r]


def box_filter(u, r):
    """u (3, n, n, n) physical -> ubar (3, nc, nc, nc), tau (6, nc, nc, nc)"""
    ubar = np.stack([_filter_sample(u[i], r) for i in range(3)])
    tau = np.stack([_filter_sample(u[i] * u[j], r) - ubar[i] * ubar[j] for i, j in SYM])
    return ubar, tau


def deviatoric(tau):
    t = np.array(tau, np.float64, copy=True)
    tr = (t[0] + t[3] + t[5]) / 3.0
    t[0] -= tr
    t[3] -= tr
    t[5] -= tr
    return t


def kwaves(g):
    """kx (num_al,1), ky (1,n), kz (num_al,1) of the cylinder entries"""
    n = g.n
    kx = g.col_x.astype(np.float64)[:, None]
    kz = np.where(g.col_z > n // 2, g.col_z - n, g.col_z).astype(np.float64)[:, None]
    y = np.arange(n)
    ky = np.where(y > n // 2, y - n, y).astype(np.float64)[None, :]
    return kx, ky, kz


def gradient(g, uy):
    """A (3, 3, n, n, n): A[i, j] = du_i/dx_j of the cylinder state uy[:3]
    (on the mask) in physical space, by i k_j through kxtran"""
    m = g.mask()
    kx, ky, kz = kwaves(g)
    ks = (kx, ky, kz)
    d = np.stack([1j * ks[j] * uy[i] * m for i in range(3) for j in range(3)])
    return g.kxtran(np.ascontiguousarray(d), 9).reshape(3, 3, g.n, g.n, g.n)


def _mm(a, b):
    return np.einsum("ik...,kj...->ij...", a, b)


def invariants(A, delta):
    """lam (5, ...) and basis T (4, 6, ...) of the gradient A (3, 3, ...)"""
    tra = (A[0, 0] + A[1, 1] + A[2, 2]) / 3.0
    S = 0.5 * (A + A.swapaxes(0, 1))
    for i in range(3):
        S[i, i] = S[i, i] - tra
    W = 0.5 * (A - A.swapaxes(0, 1))
    S2, W2 = _mm(S, S), _mm(W, W)
    s2 = S2[0, 0] + S2[1, 1] + S2[2, 2]
    w2 = W2[0, 0] + W2[1, 1] + W2[2, 2]
    q = np.maximum(s2 - w2, QEPS)
    S3 = _mm(S2, S)
    W2S = _mm(W2, S)
    W2S2 = _mm(W2, S2)
    tr = lambda M: M[0, 0] + M[1, 1] + M[2, 2]
    lam = np.stack([s2 / q, w2 / q, tr(S3) / q ** 1.5, tr(W2S) / q ** 1.5, tr(W2S2) / q ** 2])
    d2 = delta * delta
    eye = np.eye(3).reshape((3, 3) + (1,) * (A.ndim - 2))
    T = [d2 * np.sqrt(2.0 * s2) * S, d2 * (_mm(S, W) - _mm(W, S)), d2 * (S2 - s2 / 3.0 * eye),
         d2 * (W2 - w2 / 3.0 * eye)]
    T = np.stack([np.stack([t[i, j] for i, j in SYM]) for t in T])
    return lam, T


def conv_inputs(lam, pts):
    """X0 (P, 135) of the points pts (P, 3) = (y, z, x): lam[ch] at the 27
    periodic neighbours, k = ch * 27 + ((dy + 1) * 3 + dz + 1) * 3 + dx + 1"""
    n = lam.shape[1]
    y, z, x = pts[:, 0], pts[:, 1], pts[:, 2]
    cols = []
    for ch in range(NLAM):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    cols.append(lam[ch][(y + dy) % n, (z + dz) % n, (x + dx) % n])
    return np.stack(cols, axis=1)


# ---------------------------------------------------------------------------
# network
# ---------------------------------------------------------------------------
def forward(theta, X0, lam_pt, T):
    """X0 (P, 135), lam_pt (P, 5), T (P, 4, 6) -> tau_hat (P, 6), cache"""
    p = unpack(theta)
    h0 = np.tanh(X0 @ p["W0"].T + p["b0"])
    x1 = np.concatenate([lam_pt, h0], axis=1)
    h1 = np.tanh(x1 @ p["W1"].T + p["b1"])
    h2 = np.tanh(h1 @ p["W2"].T + p["b2"])
    h3 = np.tanh(h2 @ p["W3"].T + p["b3"])
    gco = h3 @ p["W4"].T + p["b4"]
    tau_hat = np.einsum("pn,pnc->pc", gco, T)
    return tau_hat, (X0, x1, h0, h1, h2, h3, gco)


def loss_and_grad(theta, X0, lam_pt, T, tau, sigma2):
    """L = sum_p sum_c CW_c (tau_hat - tau)^2 / (P sigma2) and dL/dtheta"""
    P = X0.shape[0]
    tau_hat, (X0, x1, h0, h1, h2, h3, gco) = forward(theta, X0, lam_pt, T)
    e = tau_hat - tau
    L = float(np.sum(CW * e * e) / (P * sigma2))
    p = unpack(theta)
    dtau = 2.0 * CW * e / (P * sigma2)
    dg = np.einsum("pc,pnc->pn", dtau, T)
    gr = {"W4": dg.T @ h3, "b4": dg.sum(0)}
    d3 = (dg @ p["W4"]) * (1.0 - h3 * h3)
    gr["W3"], gr["b3"] = d3.T @ h2, d3.sum(0)
    d2 = (d3 @ p["W3"]) * (1.0 - h2 * h2)
    gr["W2"], gr["b2"] = d2.T @ h1, d2.sum(0)
    d1 = (d2 @ p["W2"]) * (1.0 - h1 * h1)
    gr["W1"], gr["b1"] = d1.T @ x1, d1.sum(0)
    d0 = (d1 @ p["W1"])[:, NLAM:] * (1.0 - h0 * h0)
    gr["W0"], gr["b0"] = d0.T @ X0, d0.sum(0)
    return L, pack(gr)


def adam(theta, m, v, grad, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """one Adam step (t = 1, 2, ...) on copies; returns (theta, m, v)"""
    m = b1 * m + (1.0 - b1) * grad
    v = b2 * v + (1.0 - b2) * grad * grad
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


# ---------------------------------------------------------------------------
# minibatch index map (shared with the GPU kernels)
# ---------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def sample_points(seed, step, batch, nsnap, npts):
    """global point indices (snapshot * npts + point) of minibatch `step`"""
    base = splitmix64((seed * 0xD1B54A32D192ED03 + step) & _M64)
    return np.array([splitmix64((base + b) & _M64) % (nsnap * npts) for b in range(batch)], np.int64)


def batch_arrays(lam_s, T_s, tau_s, idx):
    """the network inputs of global points idx from snapshot arrays
    lam_s (S, 5, n, n, n), T_s (S, 4, 6, n, n, n), tau_s (S, 6, n, n, n)
    (tau deviatoric)"""
    n = lam_s.shape[2]
    npts = n ** 3
    s, p = idx // npts, idx % npts
    y, z, x = p // (n * n), (p // n) % n, p % n
    X0 = np.concatenate([conv_inputs(lam_s[si], np.array([[yy, zz, xx]]))
                         for si, yy, zz, xx in zip(s, y, z, x)])
    lam_pt = lam_s[s, :, y, z, x]
    T = T_s[s, :, :, y, z, x]
    tau = tau_s[s, :, y, z, x]
    return X0, lam_pt, T, tau


def sigma2_of(tau_s):
    """mean Frobenius norm^2 of the (deviatoric) targets tau_s (S, 6, ...):
    the loss scale"""
    t = np.asarray(tau_s, np.float64).reshape(tau_s.shape[0], 6, -1)
    return float(np.mean(np.einsum("c,scp->sp", CW, t * t)))
