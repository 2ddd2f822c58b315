"""Benchmark of the MI355X PSDNS RK2 step (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1024] [--prec 4]

Workload: forced isotropic DNS, velocity only, n^3 (default 1024^3) FP32,
RK2 with phase shifts (kshift 1), CFL pi (Courant 0.5 in the reference's
convention), FEK forcing of shells |k| <= 2 (kf_shell 3), viscosity from
kmax*eta = 1.5.  Synthetic seeded initial field (no datasets).  A "step" is
one full RK2 step: 2 x (stage_nonlinear + stage_update) + FEK forcing, all
state resident in HBM.  Prints one JSON line.

The CPU baseline times the FP64 oracle (a port of the reference algorithm)
on a bounded sample (RK2 steps at 256^3 for about 10-20 s, all host cores)
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


def pass_bytes(n, s, num_al, max_al_x):
    """Algorithmic HBM bytes per launch of each pass (external inputs and
    outputs at storage precision s; complex = 2s):
      y-inverse   3 vars: read cylinder, write z-pencil columns     2 * 3 * num_al * n * 2s
      z-inverse   3 vars: read cylinder columns, write full lines   3 * n * max_al_x * (nkept + n) * 2s
      x fused     3 spectral in, 5 spectral out (kx < max_al_x)     8 * n^2 * max_al_x * 2s
      z-forward   5 lines in, 4 cylinder columns out                (5 * n * max_al_x * n + 4 * num_al * n) * 2s
      y-fwd+upd   4 vars gathered + un read, u and un written       (4 + 3 + 3 + 3) * num_al * n * 2s
    (nkept * max_al_x = num_al)."""
    c = 2 * s
    return {
        "y_inverse": 2 * 3 * num_al * n * c,
        "z_inverse": 3 * (num_al * n + n * max_al_x * n) * c,
        "x_fused": 8 * n * n * max_al_x * c,
        "z_forward": (5 * n * max_al_x * n + 4 * num_al * n) * c,
        "y_forward_update": 13 * num_al * n * c,
    }


def cpu_baseline(n_target, budget_s=15.0):
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle as O  # the checker; timed here only as the reported CPU baseline

    O.build()
    n = 256
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
    n, prec = args.n, args.prec
    s_bytes = 4 if prec == 4 else 8
    seed = 122929298
    print(f"[bench] n={n} prec=fp{8 * s_bytes} steps={args.steps} warmup={args.warmup}", file=sys.stderr, flush=True)
    nc = args.nc
    solver = P.Solver(n, prec, viscos=1.0, nc=nc, pr=[1.0] * nc, grad=[(1.0, 0.0, 0.0)] * nc)
    solver.init_field(seed=seed, kpeak=3.0, tke=0.5)
    # viscosity from kmax eta = 1.5: eps = 2 nu D, eta = (nu^3/eps)^(1/4)
    _, _, tke0, eps_nu1 = solver.shell_stats(0)
    D = eps_nu1 / 2.0
    kmax = math.sqrt(2.0) / 3.0 * n
    viscos = math.sqrt(2.0 * D) * (1.5 / kmax) ** 2
    solver.viscos = viscos
    kf_shell = 3
    ekinf = solver.ek_shell(kf_shell + 1)
    RunK = P.Run4 if args.rk == 4 else P.Run
    run = RunK(solver, viscos, math.pi, kf_shell=kf_shell, ekinf=ekinf)
    rng = np.random.default_rng(seed)
    gsh = np.zeros(6)
    gsh[3:] = rng.random(3)

    def next_gsh():
        r = rng.random(3)
        gsh[:3] = np.mod(gsh[3:] + 0.5, 1.0)  # shifts.f kshift = 1
        gsh[3:] = r
        return gsh.copy()

    for _ in range(args.warmup):
        run.step(next_gsh())
    print("[bench] warmup done", file=sys.stderr, flush=True)
    solver.stage_times(True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        run.step(next_gsh())
    t1 = time.perf_counter()
    times = solver.stage_times(False)
    print("[bench] timed steps done", file=sys.stderr, flush=True)
    sps = (t1 - t0) / args.steps
    _, _, tke, eps = solver.shell_stats(0)

    names = ["y_inverse", "z_inverse", "x_fused", "z_forward", "y_forward_update"]
    nb = pass_bytes(n, s_bytes, solver.num_al, solver.max_al_x)
    # launches per RK2 step: two of each pass, except the y inverse of
    # stage 2, which the stage-1 y update kernel performs (same bytes there:
    # it writes yw in place of u) when there are no scalars; the scalars'
    # own transforms are not in these pass timers
    per_step = {nm: 2 for nm in names}
    per_step["y_inverse"] = 1 if nc == 0 else 2
    if args.rk == 4:
        # four substeps, none fused across substeps; the last timer is the
        # fused y pass with advanc_K, whose accumulator traffic the y-update
        # byte model does not include
        per_step = {nm: 4 for nm in names}
    passes = {}
    for i, nm in enumerate(names):
        ms = times[i] / (per_step[nm] * args.steps)
        gbps = None if args.rk == 4 and nm == "y_forward_update" else round(nb[nm] / (ms * 1e-3) / 1e9, 1)
        passes[nm] = {"ms": round(ms, 4), "GBps": gbps}
    timed = names if args.rk == 2 else names[:4]
    fft_bytes = sum(nb[nm] * per_step[nm] for nm in timed)
    fft_ms = sum(times[names.index(nm)] for nm in timed) / args.steps
    dominant = max(timed, key=lambda k: passes[k]["ms"])
    dom_ms = times[names.index(dominant)] / (per_step[dominant] * args.steps)
    achieved = nb[dominant] / (dom_ms * 1e-3) / 1e9

    # HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE counter
    # passes over this same workload (scaled by a plain copy's readings)
    traffic = None
    tfile = REPO / "profiles" / f"traffic_{n}_fp{8 * s_bytes}.json"
    if tfile.exists():
        kern = json.loads(tfile.read_text()).get("kernels", {})
        traffic = kern.get(dominant, {}).get("bytes_per_launch")

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
            "n": n, "nc": nc, "viscos": viscos, "cfl": math.pi, "kmax_eta": 1.5,
            "grid_points_per_s": n ** 3 / sps,
            "tke": tke, "epsilon": eps, "mean_dt": float(np.mean(run.dts[-args.steps:])),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_launch": nb[dominant],
            "ms_per_launch": round(dom_ms, 4),
        },
        "fft_passes": {
            "per_step_ms": round(fft_ms, 3),
            "alg_GB_per_step": round(fft_bytes / 1e9, 3),
            "frac": round(fft_bytes / (fft_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "passes": passes,
        },
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(n)
    else:
        line["cpu_baseline"] = None
    solver.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
