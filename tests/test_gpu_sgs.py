"""Learned SGS closure on the GPU (include/psdns_sgs.h) against the FP64
oracle (oracle/sgs.py): box filter and subgrid stress, gradient,
invariants and tensor basis (unshifted and on a phase-shifted grid),
FP32 and bf16-matrix-core inference, the minibatch loss and gradient of
the training kernels, the Adam update, and a small end-to-end run that
trains on box-filtered GPU DNS snapshots and beats the closure-free model
on a held-out snapshot."""
import numpy as np
import pytest

from conftest import load_oracle, rel

load_oracle()
import sgs as SG  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,nc,prec", [(64, 32, 8), (128, 32, 4)])
def test_box_filter_matches_oracle(O, P, n, nc, prec):
    print("[gpu case] sgs box filter", n, nc, prec, flush=True)
    g = O.Grid(n)
    u0 = g.field(seed=41)
    phys = g.kxtran(u0, 3)
    ubar, tau = SG.box_filter(phys, n // nc)
    tau = SG.deviatoric(tau)
    fine, coarse = P.Solver(n, prec), P.Solver(nc, prec)
    m = P.Closure(nc)
    try:
        fine.put_state(u0)
        fine.kxtran(3)
        m.filter(fine, coarse)
        got_u = coarse.get_field(P.FIELD_UX, 3)
        got_t = m.field(P.SGS_TARGET)
    finally:
        m.close()
        fine.close()
        coarse.close()
    assert rel(got_u, ubar) < (1e-12 if prec == 8 else 1e-6)
    assert rel(got_t, tau) < 2e-6  # stored in FP32
    assert np.abs(tau).max() > 1e-3


@pytest.mark.parametrize("shifted", [False, True], ids=["unshifted", "phase1"])
def test_features_match_oracle(O, P, shifted):
    print("[gpu case] sgs features", shifted, flush=True)
    n = 32
    g = O.Grid(n)
    uc = g.field(seed=43)
    gsh = np.array([0.21, 0.63, 0.37, 0.71, 0.12, 0.45])
    s = P.Solver(n, 8)
    m = P.Closure(n)
    try:
        s.put_state(uc)
        uy = uc.copy()
        if shifted:
            s.set_shifts(gsh)
            s.phase1(1)
            g.set_shifts(gsh)
            g.phshift_inplace(uy, 3, 1)
        m.features(s)
        A_got = m.field(P.SGS_GRADIENT)
        lam_got, T_got = m.field(P.SGS_LAMBDA), m.field(P.SGS_BASIS)
    finally:
        m.close()
        s.close()
    A = SG.gradient(g, uy)
    lam, T = SG.invariants(A, 2 * np.pi / n)
    assert rel(A_got, A) < 1e-6
    assert rel(lam_got, lam) < 1e-5
    assert rel(T_got, T) < 1e-5
    # incompressible: tr A vanishes to the transform round-off
    assert np.abs(A[0, 0] + A[1, 1] + A[2, 2]).max() < 1e-10 * np.abs(A).max()


def _features(O, n, seed):
    g = O.Grid(n)
    A = SG.gradient(g, g.field(seed=seed))
    lam, T = SG.invariants(A, 2 * np.pi / n)
    return lam.astype(np.float32), T.astype(np.float32)


def _theta(seed):
    th = SG.init_params(seed)
    th += 0.05 * np.random.default_rng(seed + 1).standard_normal(th.size)
    return th.astype(np.float32)


def test_predict_fp32_and_bf16_match_oracle(O, P):
    print("[gpu case] sgs predict", flush=True)
    n = 32
    lam, T = _features(O, n, 47)
    theta = _theta(3)
    m = P.Closure(n)
    try:
        m.put_fields(lam, T)
        m.set_params(theta)
        m.predict(bf16=False)
        got32 = m.field(P.SGS_PREDICTED)
        m.predict(bf16=True)
        got16 = m.field(P.SGS_PREDICTED)
    finally:
        m.close()
    idx = np.arange(n ** 3)
    pts = np.stack([idx // (n * n), (idx // n) % n, idx % n], axis=1)
    X0 = SG.conv_inputs(lam.astype(np.float64), pts)
    want, _ = SG.forward(theta.astype(np.float64), X0, lam.reshape(5, -1).T.astype(np.float64),
                         T.reshape(4, 6, -1).transpose(2, 0, 1).astype(np.float64))
    want = want.T.reshape(6, n, n, n)
    e32, e16 = rel(got32, want), rel(got16, want)
    print(f"inference vs FP64 oracle: FP32 {e32:.2e}, bf16 matrix cores {e16:.2e}", flush=True)
    assert e32 < 1e-5
    assert e16 < 3e-2


def _store(O, P, n, nsnap, seed):
    m = P.Closure(n, nsnap)
    rng = np.random.default_rng(seed)
    lam_s, T_s, tau_s = [], [], []
    for i in range(nsnap):
        lam, T = _features(O, n, seed + 10 * i)
        tau = (0.02 * rng.standard_normal((6, n, n, n))).astype(np.float32)
        m.put_fields(lam, T, tau)
        m.add_snapshot()
        lam_s.append(lam)
        T_s.append(T)
        tau_s.append(tau)
    return m, np.array(lam_s, np.float64), np.array(T_s, np.float64), np.array(tau_s, np.float64)


def test_minibatch_gradient_matches_oracle(O, P):
    print("[gpu case] sgs gradient", flush=True)
    n, nsnap, B, seed, step = 32, 2, 1024, 5, 3
    m, lam_s, T_s, tau_s = _store(O, P, n, nsnap, 61)
    theta = _theta(7)
    try:
        assert m.snapshots() == nsnap
        sigma2 = m.target_scale()
        assert abs(sigma2 - SG.sigma2_of(tau_s)) <= 1e-12 * sigma2
        m.set_params(theta)
        grad, loss = m.gradient(B, seed, step, sigma2)
    finally:
        m.close()
    idx = SG.sample_points(seed, step, B, nsnap, n ** 3)
    X0, lam_pt, T, tau = SG.batch_arrays(lam_s, T_s.reshape(nsnap, 4, 6, n, n, n), tau_s, idx)
    L, g = SG.loss_and_grad(theta.astype(np.float64), X0, lam_pt, T, tau, sigma2)
    print(f"loss {loss:.6e} vs {L:.6e}; gradient rel L2 {rel(grad, g):.2e}", flush=True)
    assert abs(loss - L) <= 1e-5 * L
    assert rel(grad, g) < 1e-4
    off = 0
    for name, shape in SG.PARAM_SHAPES:  # every block, not only the largest
        cnt = int(np.prod(shape))
        assert rel(grad[off:off + cnt], g[off:off + cnt]) < 1e-3, name
        off += cnt


def test_adam_step_matches_oracle(O, P):
    print("[gpu case] sgs adam", flush=True)
    n, nsnap, B, seed, lr = 32, 2, 1024, 9, 1e-3
    m, lam_s, T_s, tau_s = _store(O, P, n, nsnap, 71)
    theta = _theta(11)
    try:
        sigma2 = m.target_scale()
        m.set_params(theta)
        losses = m.train(2, B, lr, seed, sigma2)
        got = m.get_params()
    finally:
        m.close()
    th = theta.astype(np.float64)
    mm, vv = np.zeros_like(th), np.zeros_like(th)
    for t in (1, 2):
        idx = SG.sample_points(seed, t, B, nsnap, n ** 3)
        X0, lam_pt, T, tau = SG.batch_arrays(lam_s, T_s.reshape(nsnap, 4, 6, n, n, n), tau_s, idx)
        L, g = SG.loss_and_grad(th, X0, lam_pt, T, tau, sigma2)
        assert abs(losses[t - 1] - L) <= 1e-4 * L
        th, mm, vv = SG.adam(th, mm, vv, g, t, lr)
    assert rel(got - theta, th - theta.astype(np.float64)) < 1e-2


def test_training_on_dns_snapshots_beats_no_model(O, P):
    """128^3 FP32 forced DNS on the GPU, spun up for 400 steps (about one
    eddy turnover, so that the cascade fills the subgrid scales), then
    box-filtered to 32^3 (r = 4) every 10 steps: 5 training snapshots and a
    held-out sixth.  After training the closure's loss on the held-out
    snapshot is well below that of no model (tau_hat = 0, loss 1 in units
    of the held-out target's scale)."""
    print("[gpu case] sgs training on DNS snapshots", flush=True)
    n, nc, nu = 128, 32, 0.01
    g = O.Grid(n)
    u0 = g.field(seed=122929298)
    rng = np.random.default_rng(99)
    dns, les = P.Solver(n, 4, viscos=nu), P.Solver(nc, 4, viscos=nu)
    m = P.Closure(nc, 5)
    try:
        dns.put_state(u0)
        run = P.Run(dns, nu, np.pi, kf_shell=3)
        for _ in range(400):
            run.step(rng.random(6))
        for i in range(50):
            run.step(rng.random(6))
            if i % 10 == 9:
                P.training_snapshot(m, dns, les)
        assert m.snapshots() == 5
        sigma2 = m.target_scale()
        losses = m.train(600, 4096, 2e-3, 17, sigma2)
        # held-out snapshot ten steps later
        for _ in range(10):
            run.step(rng.random(6))
        dns.kxtran(3)
        m.filter(dns, les)
        les.xktran(3)
        m.features(les)
        sig_h = float(np.mean(np.sum(SG.CW[:, None] * m.field(P.SGS_TARGET).reshape(6, -1).astype(np.float64) ** 2, axis=0)))
        m.predict(bf16=False)
        trained = m.eval(sig_h)
        m.predict(bf16=True)
        trained16 = m.eval(sig_h)
        m.set_params(np.zeros(m.nparam, np.float32))
        m.predict()
        none = m.eval(sig_h)
    finally:
        m.close()
        dns.close()
        les.close()
    print(f"training loss {losses[0]:.3f} -> {losses[-1]:.3f}; held-out loss: no model {none:.3f}, "
          f"trained FP32 {trained:.3f}, trained bf16 {trained16:.3f}", flush=True)
    assert abs(none - 1.0) < 1e-6
    assert np.mean(losses[-20:]) < 0.7 * np.mean(losses[:20])
    assert trained < 0.7 * none
    assert trained16 < 0.75 * none


def test_les_without_closure_is_the_dns_step(O, P):
    """the per-routine LES loop with no model steps like the fused RK2 step"""
    print("[gpu case] les loop vs fused step", flush=True)
    n = 32
    g = O.Grid(n)
    u0 = g.field(seed=5)
    a, b = P.Solver(n, 8, viscos=0.02), P.Solver(n, 8, viscos=0.02)
    try:
        a.put_state(u0)
        b.put_state(u0)
        ra = P.Run(a, 0.02, np.pi, kf_shell=3)
        rb = P.Les(b, 0.02, np.pi, kf_shell=3)
        rng = np.random.default_rng(4)
        for _ in range(3):
            gsh = rng.random(6)
            ra.step(gsh)
            rb.step(gsh)
        assert rel(b.get_state(), a.get_state(), g.mask()[None]) < 1e-12
    finally:
        a.close()
        b.close()


def test_les_with_closure_adds_the_stress_divergence(O, P):
    """One LES stage with a closure whose stress is a fixed field (the
    basis coefficients constant: g = b4): the LES step minus the plain step
    equals the oracle's step of the products with tau added"""
    print("[gpu case] les stress term", flush=True)
    n, nu = 32, 0.02
    g = O.Grid(n)
    u0 = g.field(seed=6)
    gsh = np.random.default_rng(8).random(6)
    theta = np.zeros(SG.NPARAM, np.float32)
    theta[-4:] = [0.05, -0.02, 0.03, 0.01]  # g = b4 everywhere
    s = P.Solver(n, 8, viscos=nu)
    m = P.Closure(n)
    try:
        s.put_state(u0)
        m.set_params(theta)
        les = P.Les(s, nu, np.pi, closure=m, bf16=False)
        les.step(gsh)
        got = s.get_state()
    finally:
        m.close()
        s.close()
    # oracle: the same stage sequence with tau = sum_n g_n T_n(shifted grid)
    uy = g.zeros_cyl(5)
    uy[:3] = u0
    uny = u0.copy()
    g.set_shifts(gsh)
    for k in (1, 2):
        g.phshift_inplace(uy, 3, k)
        A = SG.gradient(g, uy)
        _, T = SG.invariants(A, 2 * np.pi / n)
        tau = np.einsum("n,nc...->c...", theta[-4:].astype(np.float64), T)
        ux = g.itransform_vel(uy)
        vm = g.realspace_vel(ux, k)
        ux[0] += tau[0] - tau[3]
        ux[1] += tau[1]
        ux[2] += tau[5] - tau[3]
        ux[3] += tau[4]
        ux[4] += tau[2]
        nl = g.transform(ux)
        if k == 1:
            dt = np.pi / vm
            fac = g.step_factors(nu, dt)
        g.wavespace_vel(nl, k, dt / 2.0)
        uy = nl
        if k == 1:
            g.proc4a(uy, uny, *fac)
        else:
            g.proc4b(uy, uny)
    err = rel(got, uny, g.mask()[None])
    print("LES step with a fixed stress vs oracle:", err, flush=True)
    assert err < 1e-6  # the basis passes through FP32


@pytest.mark.parametrize("prec", [8, 4])
def test_les_fused_matches_per_routine(O, P, prec):
    """The LES through the fused entry points (features on the stage's
    shifted grid, the stress added in the x pass) steps like the
    per-routine LES loop (features after phase1, the stress added to ux)"""
    print("[gpu case] fused les vs per-routine les", prec, flush=True)
    n, nu = 32, 0.02
    g = O.Grid(n)
    u0 = g.field(seed=9)
    theta = _theta(13)
    out = []
    for fused in (False, True):
        s = P.Solver(n, prec, viscos=nu)
        m = P.Closure(n)
        try:
            s.put_state(u0)
            m.set_params(theta)
            les = P.Les(s, nu, np.pi, closure=m, bf16=False, kf_shell=3, fused=fused)
            rng = np.random.default_rng(21)
            for _ in range(3):
                les.step(rng.random(6))
            out.append(s.get_state())
        finally:
            m.close()
            s.close()
    err = rel(out[1], out[0], g.mask()[None])
    print(f"fused vs per-routine LES after 3 steps: rel L2 {err:.2e}", flush=True)
    assert err < (1e-10 if prec == 8 else 2e-5)


def test_les_stage_with_scalars_is_refused(O, P):
    """The fused x pass carries either the subgrid stress or the scalars'
    advection terms: attaching a stress to a context with passive scalars
    is refused with PSDNS_EINVAL and leaves the context usable."""
    n = 32
    print("[gpu case] attach stress with nc = 1", n, flush=True)
    s = P.Solver(n, 4, viscos=0.01, nc=1, pr=[1.0], grad=[[1.0, 0.0, 0.0]])
    m = P.Closure(n)
    try:
        s.init_field(seed=5, kpeak=3.0, tke=0.5)
        m.set_params(_theta(7))
        rng = np.random.default_rng(3)
        m.put_fields(rng.standard_normal((5, n, n, n)), rng.standard_normal((24, n, n, n)))
        m.predict(True)
        with pytest.raises(P.PsdnsError, match="passive scalars"):
            m.attach(s, 1)
        s.shell_stats(0)  # the context still works
    finally:
        m.close()
        s.close()
