"""Learned-closure LES experiment (the forced-DNS-trained SGS closure on a
coarse grid), run on one GPU through the library's C ABI.

    python nasa-bluewater-psdns-benchmark-minsky_amd/les_experiment.py \
        [--n-dns 512] [--n-les 128] [--snapshots 200] [--spacing 0.1] \
        [--spin 1.0] [--eval 2.0] [--train-steps 3000] [--batch 16384] > result.json

1. Forced DNS at n_dns^3 FP32: seeded random-phase solenoidal field with
   E(k) ~ k^4 exp(-2 (k/3)^2), TKE 0.5, viscosity from kmax eta = 1.5, FEK
   forcing of shells |k| <= 2, RK2 with phase shifts, CFL pi.  The eddy
   turnover time T_e = L / u' comes from the initial spectrum (L = pi /
   (2 u'^2) sum E(k) / k, u' = sqrt(2 TKE / 3)).
2. After a spin-up of `spin` T_e, every `spacing` T_e the DNS velocity is
   box-filtered to the LES grid (and to n_les / 2) and stored as a training
   snapshot (features + target stress) on the GPU.
3. The closure is trained on the GPU (FP32, Adam) on the n_les snapshots.
4. Evaluation over `eval` T_e from the filtered DNS state at the end of the
   collection: the DNS continues; an LES with the trained closure (bf16
   matrix-core inference) and an LES without a model start from the filtered
   field.  Spectra are averaged over the last turnover (every 0.1 T_e): the
   filtered DNS spectrum against each LES (both through the fused step
   entry points, the closure's stress added in the x pass) for shells 1 .. k_c =
   floor(sqrt(2) n_les / 3) (the shells inside the LES mask).  The LES is
   accepted when its largest relative shell error is within 5%; otherwise
   the result says it falls back to the DNS.
5. Throughput: simulated time per wall-clock second of the DNS and of the
   LES (each step synchronous at return), and their ratio.
Prints one JSON object.
"""
import argparse
import importlib.util
import json
import math
import pathlib
import sys
import time

import numpy as np

PKG = pathlib.Path(__file__).resolve().parent


def load_package():
    if "psdns_amd" in sys.modules:
        return sys.modules["psdns_amd"]
    spec = importlib.util.spec_from_file_location("psdns_amd", PKG / "__init__.py", submodule_search_locations=[str(PKG)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["psdns_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[les] {msg}", file=sys.stderr, flush=True)


class Shifts:
    """kshift = 1 phase-shift offsets (shifts.f): stage 2 of a step takes
    the stage-1 offsets of the next step shifted by half a cell"""

    def __init__(self, seed):
        self.rng = np.random.default_rng(seed)
        self.g = np.zeros(6)
        self.g[3:] = self.rng.random(3)

    def __call__(self):
        r = self.rng.random(3)
        self.g[:3] = np.mod(self.g[3:] + 0.5, 1.0)
        self.g[3:] = r
        return self.g.copy()


def viscosity_for(solver, n):
    """kmax eta = 1.5 with eps = 2 nu D: nu = sqrt(2 D) (1.5 / kmax)^2"""
    _, _, _, eps_nu1 = solver.shell_stats(0)  # solver.viscos = 1
    kmax = math.sqrt(2.0) / 3.0 * n
    return math.sqrt(eps_nu1) * (1.5 / kmax) ** 2


def turnover_time(solver):
    ek, _, tke, _ = solver.shell_stats(solver.n)
    k = np.arange(1, len(ek) + 1, dtype=np.float64)
    up = math.sqrt(2.0 * tke / 3.0)
    L = math.pi / (2.0 * up * up) * float(np.sum(ek / k))
    return L / up, L, up


def filtered_state(P, m, dns, les_tmp):
    """the box-filtered DNS velocity on the coarse grid, spectral (cylinder)"""
    dns.kxtran(3)
    m.filter(dns, les_tmp)
    les_tmp.xktran(3)
    return les_tmp.get_field(P.FIELD_U, 3)


def spectrum_of(solver, state, kc):
    solver.put_state(state)
    ek, _, _, _ = solver.shell_stats(kc)
    return ek


def run(args):
    P = load_package()
    nd, nl = args.n_dns, args.n_les
    kc = int(math.floor(math.sqrt(2.0) * nl / 3.0))
    out = {"config": vars(args).copy(), "k_c": kc}
    dns = P.Solver(nd, 4, viscos=1.0)
    dns.init_field(seed=args.seed, kpeak=3.0, tke=0.5)
    nu = viscosity_for(dns, nd)
    dns.viscos = nu
    te, L, up = turnover_time(dns)
    out.update(viscos=nu, eddy_turnover=te, integral_scale=L, u_rms=up)
    log(f"DNS {nd}^3 FP32 nu={nu:.3e} T_e={te:.3f}")
    run_d = P.Run(dns, nu, math.pi, kf_shell=3)
    sh = Shifts(args.seed)
    les_a = P.Solver(nl, 4, viscos=nu)
    les_b = P.Solver(nl // 2, 4, viscos=nu)
    m = P.Closure(nl, args.snapshots)
    mb = P.Closure(nl // 2, args.snapshots)
    t = 0.0
    dns_wall, dns_sim = 0.0, 0.0

    def dns_step():
        nonlocal t, dns_wall, dns_sim
        a = time.perf_counter()
        run_d.step(sh())
        dns_wall += time.perf_counter() - a
        dns_sim += run_d.dt
        t += run_d.dt

    while t < args.spin * te:
        dns_step()
    log(f"spin-up done at t={t:.3f} ({len(run_d.dts)} steps)")
    next_t = t
    while m.snapshots() < args.snapshots:
        if t >= next_t:
            P.training_snapshot(m, dns, les_a)
            P.training_snapshot(mb, dns, les_b)
            next_t += args.spacing * te
            if m.snapshots() % 20 == 0:
                log(f"{m.snapshots()} snapshots at t={t:.3f}")
        dns_step()
    out["dns_steps_to_collect"] = len(run_d.dts)
    # training (the n_les closure; the n_les/2 store is kept for comparison losses)
    sigma2 = m.target_scale()
    theta0 = np.zeros(m.nparam, np.float32)
    rng = np.random.default_rng(args.seed)
    # Glorot-uniform initial weights, zero biases, output layer x 0.1
    shapes = [(16, 135), (16,), (64, 21), (64,), (64, 64), (64,), (64, 64), (64,), (4, 64), (4,)]
    o = 0
    for s in shapes:
        cnt = int(np.prod(s))
        if len(s) == 2:
            a = math.sqrt(6.0 / (s[0] + s[1])) * (0.1 if s[0] == 4 else 1.0)
            theta0[o:o + cnt] = rng.uniform(-a, a, cnt)
        o += cnt
    m.set_params(theta0)
    a = time.perf_counter()
    n1 = int(args.train_steps * 0.7)
    losses = list(m.train(n1, args.batch, args.lr, args.seed, sigma2))
    losses += list(m.train(args.train_steps - n1, args.batch, args.lr * 0.3, args.seed + 1, sigma2))
    train_s = time.perf_counter() - a
    out["training"] = {"snapshots": m.snapshots(), "steps": args.train_steps, "batch": args.batch,
                       "seconds": train_s, "loss_first": float(np.mean(losses[:20])),
                       "loss_last": float(np.mean(losses[-20:])), "sigma2": sigma2}
    log(f"trained: loss {out['training']['loss_first']:.3f} -> {out['training']['loss_last']:.3f} in {train_s:.1f} s")
    theta = m.get_params()
    mb.set_params(theta)  # the same closure on the n_les/2 grid (delta scales with the grid)
    # a-priori on a held-out filtered field (after the collection window)
    dns_step()
    dns.kxtran(3)
    m.filter(dns, les_a)
    les_a.xktran(3)
    m.features(les_a)
    sig_h = float(np.mean(np.sum(np.array([1, 2, 2, 1, 2, 1.0])[:, None] *
                                 m.field(P.SGS_TARGET).reshape(6, -1).astype(np.float64) ** 2, axis=0)))
    apriori = {}
    for name, bf in (("fp32", False), ("bf16", True)):
        m.predict(bf)
        apriori[name] = m.eval(sig_h)
    m.set_params(np.zeros(m.nparam, np.float32))
    m.predict()
    apriori["no_model"] = m.eval(sig_h)
    m.set_params(theta)
    out["apriori_heldout_loss"] = apriori
    log(f"a-priori held-out loss {apriori}")

    # a-posteriori: LES with and without the closure from the filtered field
    u0 = filtered_state(P, m, dns, les_a)
    spare = P.Solver(nl, 4, viscos=nu)
    les_m = P.Solver(nl, 4, viscos=nu)
    les_n = P.Solver(nl, 4, viscos=nu)
    les_m.put_state(u0)
    les_n.put_state(u0)
    les_m_run = P.Les(les_m, nu, args.les_cfl, closure=m, bf16=True, kf_shell=3, ekinf=None, fused=True)
    les_n_run = P.Les(les_n, nu, args.les_cfl, closure=None, kf_shell=3, ekinf=les_m_run.ekinf, fused=True)
    runs = {"model": (les_m_run, les_m), "none": (les_n_run, les_n)}
    # further output gains: the closure with its linear output layer (the
    # tensor-basis coefficients) scaled by c, so that the stress is c times
    # the trained prediction; each in its own LES beside the same DNS
    gain_closures = []
    nw = sum(int(np.prod(sh)) for sh in shapes[:-2])
    for c in args.gains[1:]:
        mg = P.Closure(nl)
        th = theta.copy()
        th[nw:] *= c
        mg.set_params(th)
        sg = P.Solver(nl, 4, viscos=nu)
        sg.put_state(u0)
        runs[f"gain_{c:g}"] = (P.Les(sg, nu, args.les_cfl, closure=mg, bf16=True, kf_shell=3, ekinf=les_m_run.ekinf,
                                     fused=True), sg)
        gain_closures.append(mg)
    if args.gains[0] != 1.0:
        th = theta.copy()
        th[nw:] *= args.gains[0]
        m.set_params(th)
    t0 = t
    t_les = {k: t0 for k in runs}
    wall = {k: 0.0 for k in runs}
    sims = {k: 0.0 for k in runs}
    shf = {k: Shifts(args.seed + 7) for k in runs}
    acc = {"dns": []}
    acc.update({k: [] for k in runs})
    next_sample = t0 + (args.eval - 1.0) * te
    end = t0 + args.eval * te
    while t < end:
        dns_step()
        for key, (rr, s_) in runs.items():
            while t_les[key] < t:
                a = time.perf_counter()
                rr.step(shf[key]())
                wall[key] += time.perf_counter() - a
                sims[key] += rr.dt
                t_les[key] += rr.dt
        if t >= next_sample:
            acc["dns"].append(spectrum_of(spare, filtered_state(P, m, dns, les_a), kc))
            for key, (rr, s_) in runs.items():
                acc[key].append(s_.shell_stats(kc)[0])
            next_sample += 0.1 * te
    ed = np.mean(acc["dns"], axis=0)
    res = {}
    for key, (rr, s_) in runs.items():
        el = np.mean(acc[key], axis=0)
        err = np.abs(el - ed) / np.where(ed > 0, ed, 1.0)  # shell 1 (k = 0) holds no energy
        res[key] = {"max_rel_shell_error": float(err[1:].max()), "shell_of_max": int(np.argmax(err[1:]) + 2),
                    "rel_error_by_shell": [round(float(e), 4) for e in err],
                    "sim_time_per_wall_s": sims[key] / wall[key], "ms_per_step": 1e3 * wall[key] / len(rr.dts)}
        if key.startswith("gain_"):
            res[key]["spectrum"] = [float(v) for v in el]
    out["output_gain"] = args.gains[0]
    dns_rate = dns_sim / dns_wall
    out["dns"] = {"ms_per_step": 1e3 * dns_wall / len(run_d.dts), "sim_time_per_wall_s": dns_rate,
                  "mean_dt": dns_sim / len(run_d.dts)}
    out["les"] = res
    out["spectra"] = {"filtered_dns": [float(v) for v in ed],
                      "les_model": [float(v) for v in np.mean(acc["model"], axis=0)],
                      "les_no_model": [float(v) for v in np.mean(acc["none"], axis=0)],
                      "samples": len(acc["dns"])}
    out["speedup_vs_dns"] = res["model"]["sim_time_per_wall_s"] / dns_rate
    out["accepted"] = res["model"]["max_rel_shell_error"] <= 0.05
    out["result"] = "LES with the learned closure" if out["accepted"] else "falls back to the DNS (spectral error > 5%)"
    for s in [dns, les_a, les_b, spare] + [s_ for rr, s_ in runs.values()]:
        s.close()
    for mg in gain_closures:
        mg.close()
    m.close()
    mb.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-dns", type=int, default=512)
    ap.add_argument("--n-les", type=int, default=128)
    ap.add_argument("--snapshots", type=int, default=200)
    ap.add_argument("--spacing", type=float, default=0.1, help="eddy turnovers between snapshots")
    ap.add_argument("--spin", type=float, default=1.0, help="eddy turnovers before the first snapshot")
    ap.add_argument("--eval", type=float, default=2.0, help="eddy turnovers of the LES comparison")
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=122929298)
    ap.add_argument("--les-cfl", type=float, default=math.pi,
                    help="CFL number of the LES steps (the DNS keeps pi)")
    ap.add_argument("--gains", type=lambda v: [float(x) for x in v.split(",")], default=[1.0],
                    help="output gains c of the closure (stress = c x prediction): the first is the reported "
                         "model, each further one runs its own LES beside the same DNS")
    args = ap.parse_args()
    print(json.dumps(run(args)), flush=True)


if __name__ == "__main__":
    main()
