"""Benchmark of the MI355X PSDNS RK2 step (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1024] [--prec 4]

Workload: forced isotropic DNS, velocity only, n^3 (default 1024^3) FP32,
RK2 with phase shifts (kshift 1), CFL pi (Courant 0.5 in the reference's
convention), FEK forcing of shells |k| <= 2 (kf_shell 3), viscosity from
kmax*eta = 1.5.  Synthetic seeded initial field (no datasets).  A "step" is
one full RK2 step: 2 x (stage_nonlinear + stage_update) + FEK forcing, all
state resident in HBM.  Prints one JSON line.

After a velocity-only RK2 run the same grid is run again with two passive
scalars (Sc = 1, unit mean gradient along x, zero initial fluctuation):
its step time, per-pass device times per RK substep and the roofline of the
scalar path's dominant kernel go into the line's "config3_nc2" object
(--no-config3 skips it).  Step times are also reported as min/mean/max over
steps 2..K.  The "les_closure" object times the learned-closure LES at 128^3
against the 512^3 DNS it stands for (--no-les skips it), and
"config4_particles" a 512^3 step with 2^20 Lagrangian particles against the
same step without them (--no-particles skips it).

The CPU baseline times the FP64 oracle (a port of the reference algorithm)
on a bounded sample (RK2 steps at 512^3 for about 15-25 s, all host cores)
and scales it to n^3 by n^3 log2 n.
"""
import argparse
import importlib.util
import json
import math
import os
import pathlib
import sys
import time

import numpy as np

REPO = pathlib.Path(__file__).resolve().parent
PKG_DIR = REPO / "nasa-bluewater-psdns-benchmark-minsky_amd"
HBM_PEAK_GBS = 8000.0


def load_package():
    spec = importlib.util.spec_from_file_location("psdns_amd", PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["psdns_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pass_bytes(n, s, num_al, max_al_x, nc=0):
    """Algorithmic HBM bytes per launch of each pass (external inputs and
    outputs at storage precision s; complex = 2s):
      y-inverse   3 vars: read cylinder, write z-pencil columns     2 * 3 * num_al * n * 2s
      z-inverse   3 vars: read cylinder columns, write full lines   3 * n * max_al_x * (nkept + n) * 2s
      x fused     3 spectral in, 5 spectral out (kx < max_al_x),
                  and per scalar 3 gradient slots in, 1 out         (8 + 4 nc) * n^2 * max_al_x * 2s
      z-forward   5 lines in, 4 cylinder columns out                (5 * n * max_al_x * n + 4 * num_al * n) * 2s
      y-fwd+upd   4 vars gathered + un read; stage 1 writes un and
                  stage 2's yw, stage 2 writes un only (u = un after
                  proc4b is not stored): per launch on average
                                                                    (4 + 3 + (6 + 3) / 2) * num_al * n * 2s
    and per passive scalar:
      gradient y  1 cylinder in, the scalar and its y derivative
                  y-transformed out                                 3 * num_al * n * 2s
      z-inverse   those 2 cylinder slots in, the 3 derivatives out
                  (i kx and i kz applied in the pass)               (2 * num_al * n + 3 * n * max_al_x * n) * 2s
      x pass      inside the velocity's x fused pass                0
      z-forward   1 line set in, 1 cylinder out                      (n * max_al_x * n + num_al * n) * 2s
      y-forward   inside the update                                 0
      update      the term gathered (y forward transform) and un
                  in, u and un out                                  4 * num_al * n * 2s
    (nkept * max_al_x = num_al)."""
    c = 2 * s
    return {
        "y_inverse": 2 * 3 * num_al * n * c,
        "z_inverse": 3 * (num_al * n + n * max_al_x * n) * c,
        "x_fused": (8 + 4 * nc) * n * n * max_al_x * c,
        "z_forward": (5 * n * max_al_x * n + 4 * num_al * n) * c,
        "y_forward_update": (13 + 10) * num_al * n * c // 2,
        "sc_grad_y": 3 * num_al * n * c,
        "sc_z_inverse": (2 * num_al * n + 3 * n * max_al_x * n) * c,
        "sc_x_pass": 0,
        "sc_z_forward": (n * max_al_x * n + num_al * n) * c,
        "sc_y_forward": 0,
        "sc_update": 4 * num_al * n * c,
    }


VEL = ["y_inverse", "z_inverse", "x_fused", "z_forward", "y_forward_update"]
SC = ["sc_grad_y", "sc_z_inverse", "sc_x_pass", "sc_z_forward", "sc_y_forward", "sc_update"]


def launches_per_step(nc, rk):
    """launches of each pass per step: RK2 has two stages (the stage-2 y
    inverse of a velocity-only run is done by the stage-1 update kernel),
    RK4 four substeps; every scalar pass runs once per scalar and stage,
    except the x pass, which the velocity's x fused pass does for all, and
    the y forward pass, which is part of the scalar's update"""
    st = 2 if rk == 2 else 4
    per = {nm: st for nm in VEL}
    if rk == 2 and nc == 0:
        per["y_inverse"] = 1
    for nm in SC:
        per[nm] = st * nc if nm not in ("sc_x_pass", "sc_y_forward") else 0
    return per


def run_case(P, n, prec, nc, rk, steps, warmup, seed=122929298):
    """One timed run: seeded field, viscosity from kmax eta = 1.5, FEK
    forcing of shells |k| <= 2, kshift 1 shifts; returns timings and the
    per-pass device times (HIP events on the library's stream)."""
    s_bytes = 4 if prec == 4 else 8
    solver = P.Solver(n, prec, viscos=1.0, nc=nc, pr=[1.0] * nc, grad=[(1.0, 0.0, 0.0)] * nc)
    try:
        solver.init_field(seed=seed, kpeak=3.0, tke=0.5)
        # viscosity from kmax eta = 1.5: eps = 2 nu D, eta = (nu^3/eps)^(1/4)
        _, _, tke0, eps_nu1 = solver.shell_stats(0)
        D = eps_nu1 / 2.0
        kmax = math.sqrt(2.0) / 3.0 * n
        viscos = math.sqrt(2.0 * D) * (1.5 / kmax) ** 2
        solver.viscos = viscos
        kf_shell = 3
        ekinf = solver.ek_shell(kf_shell + 1)
        RunK = P.Run4 if rk == 4 else P.Run
        run = RunK(solver, viscos, math.pi, kf_shell=kf_shell, ekinf=ekinf)
        rng = np.random.default_rng(seed)
        gsh = np.zeros(6)
        gsh[3:] = rng.random(3)

        def next_gsh():
            r = rng.random(3)
            gsh[:3] = np.mod(gsh[3:] + 0.5, 1.0)  # shifts.f kshift = 1
            gsh[3:] = r
            return gsh.copy()

        for _ in range(warmup):
            run.step(next_gsh())
        print(f"[bench] n={n} nc={nc} rk={rk}: warmup done", file=sys.stderr, flush=True)
        solver.pass_times(True)
        step_s = []
        t0 = time.perf_counter()
        for _ in range(steps):
            a = time.perf_counter()
            run.step(next_gsh())  # every entry point is synchronous at return
            step_s.append(time.perf_counter() - a)
        t1 = time.perf_counter()
        tab = solver.pass_times(False)  # [substep][pass] ms over the timed steps
        print(f"[bench] n={n} nc={nc} rk={rk}: timed steps done", file=sys.stderr, flush=True)
        _, _, tke, eps = solver.shell_stats(0)
        num_al, max_al_x = solver.num_al, solver.max_al_x
        mean_dt = float(np.mean(run.dts[-steps:]))
    finally:
        solver.close()
    sps = (t1 - t0) / steps
    nb = pass_bytes(n, s_bytes, num_al, max_al_x, nc)
    per = launches_per_step(nc, rk)
    names = VEL + SC
    passes = {}
    for i, nm in enumerate(names):
        if per[nm] == 0:
            continue
        ms_step = float(tab[:, i].sum()) / steps
        ms = ms_step / per[nm]
        bytes_ok = not (rk == 4 and nm == "y_forward_update")  # RK4 adds the accumulator u1
        passes[nm] = {"ms": round(ms, 4), "ms_per_step": round(ms_step, 3), "launches_per_step": per[nm],
                      "GBps": round(nb[nm] / (ms * 1e-3) / 1e9, 1) if bytes_ok and ms > 0 else None}
    rest = step_s[1:] if len(step_s) > 1 else step_s
    return {
        "n": n, "nc": nc, "rk": rk, "prec": prec, "s_bytes": s_bytes, "sps": sps, "viscos": viscos,
        "tke": tke, "eps": eps, "mean_dt": mean_dt, "nb": nb, "per": per, "passes": passes, "tab": tab,
        "step_ms": {"min": round(min(rest) * 1e3, 3), "mean": round(float(np.mean(rest)) * 1e3, 3),
                    "max": round(max(rest) * 1e3, 3), "over": f"steps 2..{steps}"},
        "by_substep_ms": [[round(float(v) / steps, 3) for v in row] for row in tab[:(2 if rk == 2 else 4)]],
    }


def les_case(P, steps=10, seed=122929298):
    """The learned-closure LES against the DNS it replaces: a forced 512^3
    FP32 DNS step, and a 128^3 FP32 LES step (fused entry points, closure
    with bf16 matrix-core inference, Glorot-initialised weights: the cost
    does not depend on their values) from the DNS field box-filtered to
    128^3.  Returns the step times, the mean dt of each and the ratio of
    simulated time per wall second."""
    dns = P.Solver(512, 4, viscos=1.0)
    les = P.Solver(128, 4, viscos=1.0)
    m = P.Closure(128)
    try:
        dns.init_field(seed=seed, kpeak=3.0, tke=0.5)
        _, _, _, e1 = dns.shell_stats(0)
        nu = math.sqrt(e1) * (1.5 / (math.sqrt(2.0) / 3.0 * 512)) ** 2
        dns.viscos = les.viscos = nu
        rng = np.random.default_rng(seed)
        run = P.Run(dns, nu, math.pi, kf_shell=3)
        for _ in range(2):
            run.step(rng.random(6))
        a = time.perf_counter()
        for _ in range(steps):
            run.step(rng.random(6))
        dns_ms = (time.perf_counter() - a) / steps * 1e3
        dns_dt = float(np.mean(run.dts[-steps:]))
        th = np.zeros(m.nparam, np.float32)
        o = 0
        for shape in ((16, 135), (16,), (64, 21), (64,), (64, 64), (64,), (64, 64), (64,), (4, 64), (4,)):
            cnt = int(np.prod(shape))
            if len(shape) == 2:
                lim = math.sqrt(6.0 / sum(shape)) * (0.1 if shape[0] == 4 else 1.0)
                th[o:o + cnt] = rng.uniform(-lim, lim, cnt)
            o += cnt
        m.set_params(th)
        dns.kxtran(3)
        m.filter(dns, les)
        les.xktran(3)
        les.put_state(les.get_field(P.FIELD_U, 3))
        lr = P.Les(les, nu, math.pi, closure=m, bf16=True, kf_shell=3, fused=True)
        for _ in range(2):
            lr.step(rng.random(6))
        a = time.perf_counter()
        for _ in range(2 * steps):
            lr.step(rng.random(6))
        les_ms = (time.perf_counter() - a) / (2 * steps) * 1e3
        les_dt = float(np.mean(lr.dts[-2 * steps:]))
    finally:
        m.close()
        les.close()
        dns.close()
    return {
        "workload": "forced DNS 512^3 FP32 vs a 128^3 FP32 LES of its box-filtered field with the learned "
                    "SGS closure (bf16 matrix-core inference), both RK2 on the fused entry points",
        "dns_ms_per_step": round(dns_ms, 3), "dns_mean_dt": dns_dt,
        "les_ms_per_step": round(les_ms, 3), "les_mean_dt": les_dt,
        "speedup_vs_dns": round((les_dt / les_ms) / (dns_dt / dns_ms), 2),
        "accuracy": "spectral check of a trained closure over 2 eddy turnovers: profiles/les_config5.md "
                    "(les_experiment.py)",
    }


def particles_case(P, steps=10, seed=122929298, npart=1 << 20):
    """Configuration 4: a forced 512^3 FP32 RK2 step with 2^20 Lagrangian
    particles uniform in the box (positions 1 + U n in grid units, seed
    122929299), in partsp.f's order (rk_advpos 1, spline + interpolation at
    the new shifts, rk_advpos 2), against the same step without particles;
    then each particle entry point timed on its own (each returns after its
    kernels finish).  The interpolation gathers 4 x 4 x 4 spline
    coefficients of 3 components per particle and reads and writes 3 FP64
    position / velocity entries: 64 * 3 * 4 + 48 bytes per particle in
    FP32."""
    n = 512
    s = P.Solver(n, 4, viscos=1.0)
    try:
        s.init_field(seed=seed, kpeak=3.0, tke=0.5)
        _, _, _, e1 = s.shell_stats(0)
        nu = math.sqrt(e1) * (1.5 / (math.sqrt(2.0) / 3.0 * n)) ** 2
        s.viscos = nu
        rng = np.random.default_rng(seed)

        def timed(run, k):
            for _ in range(2):
                run.step(rng.random(6))
            a = time.perf_counter()
            for _ in range(k):
                run.step(rng.random(6))
            return (time.perf_counter() - a) / k * 1e3

        plain_ms = timed(P.Run(s, nu, math.pi, kf_shell=3), steps)
        pp = np.zeros((9, npart))
        pp[:3] = 1.0 + np.random.default_rng(122929299).random((3, npart)) * n
        run = P.Run(s, nu, math.pi, kf_shell=3, particles=pp)
        with_ms = timed(run, steps)
        dt = run.dts[-1]
        calls = {"advpos_1": lambda: s.lag_advpos(1, dt, 0.5, 0.5), "spline": s.lag_spline,
                 "interp": s.lag_interp, "advpos_2": lambda: s.lag_advpos(2, dt, 0.0, 0.5)}
        call_ms = {k: [] for k in calls}
        for _ in range(steps):
            for k, f in calls.items():
                a = time.perf_counter()
                f()
                call_ms[k].append((time.perf_counter() - a) * 1e3)
    finally:
        s.close()
    ms = {k: round(float(np.median(v)), 4) for k, v in call_ms.items()}
    interp_bytes = npart * (64 * 3 * 4 + 48)
    gbs = interp_bytes / (ms["interp"] * 1e-3) / 1e9
    return {
        "workload": "forced DNS 512^3 FP32 RK2 with 2^20 particles uniform in the box (LAG_LOCAL cubic-spline "
                    "path, partsp RK2 order), against the same step without particles",
        "particles": npart, "ms_per_step": round(with_ms, 3), "ms_per_step_no_particles": round(plain_ms, 3),
        "particle_ms_per_step": round(with_ms - plain_ms, 3), "call_ms_median": ms,
        "interp_roofline": {"bound": "hbm (gather)", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_call": interp_bytes,
                            "note": "host-timed call (kernel + launch + sync)"},
    }


def fft_summary(r, names):
    nb, per, passes = r["nb"], r["per"], r["passes"]
    timed = [nm for nm in names if nm in passes and passes[nm]["GBps"] is not None]
    fft_bytes = sum(nb[nm] * per[nm] for nm in timed)
    fft_ms = sum(passes[nm]["ms_per_step"] for nm in timed)
    return timed, fft_bytes, fft_ms


def cpu_baseline(n_target, budget_s=15.0, n=512):
    """The FP64 oracle's RK2 step on the host cores at n^3 (512^3: about
    4.5 s per step on 16 threads), timed for about budget_s and scaled to
    n_target^3 by n^3 log2 n (x8.9 from 512^3 to 1024^3; measured from 256^3
    the same scaling overestimates the 512^3 step by 15%)."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle as O  # the checker; timed here only as the reported CPU baseline

    O.build()
    n = min(n, n_target)
    g = O.Grid(n)
    u0 = g.field()
    r = O.Run(g, u0, 0.01, cfl=math.pi, kshift=1, kf_shell=3)
    r.step()  # first-touch warm-up
    t0 = time.perf_counter()
    k = 0
    while True:
        r.step()
        k += 1
        if time.perf_counter() - t0 > budget_s or k >= 20:
            break
    per = (time.perf_counter() - t0) / k
    scale = (n_target / n) ** 3 * math.log2(n_target) / math.log2(n)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": per * scale,
        "unit": "s/step",
        "cores": threads,
        "kind": "port",
        "sample": f"{k} RK2 steps of the FP64 oracle at {n}^3 ({per:.3f} s/step, {threads} OpenMP threads), "
                  f"scaled to {n_target}^3 by n^3 log2 n (x{scale:.1f})",
    }


def copy_calibration(nbytes):
    """A plain device-to-device copy of nbytes (the runtime's copy kernel), for
    scaling the FETCH_SIZE / WRITE_SIZE counters of the pass kernels."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--prec", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nc", type=int, default=0,
                    help="passive scalars (Sc = 1, unit mean gradient along x, zero initial fluctuation)")
    ap.add_argument("--rk", type=int, default=2, choices=(2, 4),
                    help="time scheme: 2 (proc4a/proc4b) or 4 (advanc_1..4)")
    ap.add_argument("--no-config3", dest="config3", action="store_false",
                    help="skip the two-scalar run that follows a velocity-only RK2 run")
    ap.add_argument("--config3-steps", type=int, default=5)
    ap.add_argument("--no-les", dest="les", action="store_false",
                    help="skip the learned-closure LES vs DNS timing")
    ap.add_argument("--no-particles", dest="particles", action="store_false",
                    help="skip the 512^3 step with 2^20 Lagrangian particles")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is a design, not an implementation)", file=sys.stderr)
        sys.exit(2)

    P = load_package()
    n, prec, nc = args.n, args.prec, args.nc
    print(f"[bench] n={n} prec=fp{8 * (4 if prec == 4 else 8)} steps={args.steps} warmup={args.warmup}",
          file=sys.stderr, flush=True)
    r = run_case(P, n, prec, nc, args.rk, args.steps, args.warmup)
    sps, nb, passes = r["sps"], r["nb"], r["passes"]
    timed, fft_bytes, fft_ms = fft_summary(r, VEL + SC)
    dominant = max(timed, key=lambda k: passes[k]["ms_per_step"])
    dom_ms = passes[dominant]["ms"]
    achieved = nb[dominant] / (dom_ms * 1e-3) / 1e9

    # HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE counter
    # passes over this same workload (scaled by a plain copy's readings)
    def traffic_of(kernel, n, s_bytes, nc):
        tfile = REPO / "profiles" / f"traffic_{n}_fp{8 * s_bytes}{'_nc%d' % nc if nc else ''}.json"
        if tfile.exists():
            return json.loads(tfile.read_text()).get("kernels", {}).get(kernel, {}).get("bytes_per_launch")
        return None

    line = {
        "metric": ("s/RK2 step & FFT HBM-roofline % at 1024^3 FP32; LES-closure speedup vs DNS" if args.rk == 2
                   else "s/RK4 step at 1024^3 FP32"),
        "value": sps,
        "unit": "s/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sps * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == 4 else "f64",
        "data": "synthetic: seeded random-phase solenoidal field, E(k) ~ k^4 exp(-2(k/3)^2), TKE 0.5",
        "config": {
            "workload": f"forced isotropic DNS {n}^3 {'FP32' if prec == 4 else 'FP64'}, "
                        + ("velocity only (nc=0), " if nc == 0 else
                           f"{nc} passive scalars (Sc=1, mean gradient along x, zero initial fluctuation), ")
                        + f"RK{args.rk} + phase shifts (kshift 1), FEK forcing kf_shell 3, CFL pi",
            "n": n, "nc": nc, "viscos": r["viscos"], "cfl": math.pi, "kmax_eta": 1.5,
            "grid_points_per_s": n ** 3 / sps,
            "tke": r["tke"], "epsilon": r["eps"], "mean_dt": r["mean_dt"],
        },
        "step_ms": r["step_ms"],
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_of(dominant, n, r["s_bytes"], nc),
            "bytes_per_launch": nb[dominant],
            "ms_per_launch": round(dom_ms, 4),
        },
        "fft_passes": {
            "per_step_ms": round(fft_ms, 3),
            "alg_GB_per_step": round(fft_bytes / 1e9, 3),
            "frac": round(fft_bytes / (fft_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "passes": passes,
            "by_substep_ms": r["by_substep_ms"],
        },
    }
    if args.config3 and nc == 0 and args.rk == 2:
        # the configuration with two passive scalars (Sc = 1, unit mean
        # gradient along x, zero initial fluctuation), same grid and precision
        r3 = run_case(P, n, prec, 2, 2, args.config3_steps, args.warmup)
        t3, b3, ms3 = fft_summary(r3, VEL + SC)
        sc_dom = max((k for k in SC if k in r3["passes"]), key=lambda k: r3["passes"][k]["ms_per_step"])
        sd_ms = r3["passes"][sc_dom]["ms"]
        sd = r3["nb"][sc_dom] / (sd_ms * 1e-3) / 1e9
        line["config3_nc2"] = {
            "workload": f"forced isotropic DNS {n}^3 {'FP32' if prec == 4 else 'FP64'} with 2 passive scalars "
                        "(Sc=1, mean gradient along x, zero initial fluctuation), RK2",
            "s_per_step": r3["sps"], "ms_per_step": round(r3["sps"] * 1e3, 3), "steps": args.config3_steps,
            "step_ms": r3["step_ms"],
            "grid_points_per_s": n ** 3 / r3["sps"],
            "fft_passes": {"per_step_ms": round(ms3, 3), "alg_GB_per_step": round(b3 / 1e9, 3),
                           "frac": round(b3 / (ms3 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                           "passes": r3["passes"], "by_substep_ms": r3["by_substep_ms"]},
            "scalar_roofline": {"bound": "hbm", "kernel": sc_dom, "achieved": round(sd, 1), "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "frac": round(sd / HBM_PEAK_GBS, 4),
                                "traffic": traffic_of(sc_dom, n, r3["s_bytes"], 2),
                                "bytes_per_launch": r3["nb"][sc_dom], "ms_per_launch": round(sd_ms, 4)},
        }
    if args.les and args.rk == 2 and n == 1024:
        print("[bench] les vs dns", file=sys.stderr, flush=True)
        line["les_closure"] = les_case(P)
    if args.particles and args.rk == 2 and n == 1024:
        print("[bench] particles", file=sys.stderr, flush=True)
        line["config4_particles"] = particles_case(P)
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(n)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
