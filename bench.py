"""Headline benchmark: time-to-solution at 1e-10 true residual on a 32^3x64
clover lattice (gauge eps 0.2, csw 1.0, kappa 0.1248, antiperiodic t), fast
path (learned preconditioner when trained, mixed FP16-with-scale inner
BiCGSTAB with FP64 reliable updates) against the plain FP64 even-odd BiCGSTAB
on the same GPU.

A step is one complete solve of the fast path on one Gaussian source that is
already resident in HBM (external PyQuda layout).  The plain FP64 solve of the
same source is timed alongside; value = t_plain / t_fast.

Prints one JSON line (rank 0, N = 1).  --gpus N > 1 exits: this build runs on
one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-solution @1e-10 resid, 32^3x64 clover: learned+mixed vs plain FP64"
L = (32, 32, 32, 64)
EPS, CSW, KAPPA, TB = 0.2, 1.0, 0.1248, -1
TOL = 1e-10
PEAK_GBS = 8000.0
# algorithmic bytes per output site of one hop launch (8 neighbour spinors,
# 8 compressed links, 1 output) and of the packed clover inverse, per precision
HOP_BYTES = {0: 8 * 192 + 8 * 96 + 192, 1: 8 * 96 + 8 * 48 + 96, 2: 8 * 52 + 8 * 24 + 52}
CLOVER_BYTES = {0: 576, 1: 288, 2: 148}
# the xpay epilogue of the Schur operator reads one more spinor
XPAY_BYTES = {0: 192, 1: 96, 2: 52}
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "hop_fp16_clover_32x32x32x64_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(U_host, b_host, iters_gpu_plain, sample_iters, threads):
    """FP64 OpenMP host oracle on a bounded sample: `sample_iters` BiCGSTAB
    iterations of the same system, extrapolated to the iteration count of the
    GPU's plain FP64 solve."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.build()
    if threads:
        oracle.set_threads(threads)
    nthreads = oracle.threads()
    _, info = oracle.bicgstab(b_host, U_host, L, KAPPA, CSW, TB, tol=1e-30, max_iter=sample_iters)
    per_iter = info["iter_s"] / max(info["iters"], 1)
    est = info["setup_s"] + per_iter * iters_gpu_plain
    return {"value": round(est, 3), "unit": "s (FP64 time-to-solution, extrapolated)",
            "cores": nthreads, "kind": "port",
            "sample": f"{info['iters']} BiCGSTAB iterations of the same 32^3x64 clover system "
                      f"({info['iter_s']:.1f} s) + clover set-up ({info['setup_s']:.1f} s), "
                      f"extrapolated to {iters_gpu_plain} iterations"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-iters", type=int, default=6, help="oracle iterations in the CPU sample")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--inner", default="fp16", choices=["fp32", "fp16"])
    ap.add_argument("--delta", type=float, default=0.1)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU halo exchange is out of scope)", file=sys.stderr)
        sys.exit(2)

    import torch
    from qcu_amd.lattice import FP16, FP32, FP64, Lattice, gaussian, random_gauge

    inner = FP16 if args.inner == "fp16" else FP32
    U = random_gauge(L, EPS, 20241)
    lat = Lattice(L, KAPPA, CSW, TB)
    lat.load_gauge(U)
    b = gaussian((2,) + (L[3], L[2], L[1], L[0] // 2, 4, 3), 777)
    torch.cuda.synchronize()
    log("gauge + clover resident; tuning hop launch shapes")
    for p in (FP64, inner):
        log(f"  prec {p}: {lat.tune(p)}")

    def fast():
        return lat.solve(b, tol=TOL, prec=inner, reliable_delta=args.delta)

    def plain():
        return lat.solve(b, tol=TOL, prec=FP64)

    for _ in range(args.warmup):
        fast()
    x, r = plain()
    log(f"plain warmup: {r}")
    torch.cuda.synchronize()
    t_plain = []
    plain_res = None
    for _ in range(max(1, args.steps)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, plain_res = plain()
        torch.cuda.synchronize()
        t_plain.append(time.perf_counter() - t0)
    log(f"plain: {plain_res}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fast_res = []
    for _ in range(args.steps):
        x, r = fast()
        fast_res.append(r)
    torch.cuda.synchronize()
    t_fast = (time.perf_counter() - t0) / args.steps
    log(f"fast: {fast_res[-1]}")
    t_plain_med = sorted(t_plain)[len(t_plain) // 2]
    assert all(r.converged and r.true_res <= TOL for r in fast_res), "fast solve did not reach 1e-10"
    assert plain_res.converged and plain_res.true_res <= TOL, "plain solve did not reach 1e-10"

    # roofline of the dominant kernel: the inner-precision clover hop with the
    # xpay epilogue, timed with HIP events on the library's stream
    vh = lat.vh
    ms_hop = lat.bench_hop(prec=inner, clover=True, xpay=True, iters=100)
    alg = (HOP_BYTES[inner] + CLOVER_BYTES[inner] + XPAY_BYTES[inner]) * vh
    achieved = alg / (ms_hop * 1e-3) / 1e9
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    ms64 = lat.bench_hop(prec=FP64, clover=True, xpay=True, iters=50)
    alg64 = (HOP_BYTES[0] + CLOVER_BYTES[0] + XPAY_BYTES[0]) * vh

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle, bounded sample)")
        cpu = cpu_baseline(U.cpu().numpy().reshape(4, 2, vh, 3, 3), b.cpu().numpy().reshape(2, vh, 4, 3),
                           plain_res.iters, args.cpu_iters, args.cpu_threads)

    out = {
        "metric": METRIC,
        "value": round(t_plain_med / t_fast, 3),
        "unit": "x (plain FP64 time / fast time)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_fast * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (outer) / " + args.inner + " (inner)",
        "data": "synthetic: U = exp(i 0.2 H) seeded, Gaussian source",
        "config": {"workload": "32^3x64 clover, eps 0.2, csw 1.0, kappa 0.1248, antiperiodic t, "
                               "tol 1e-10 true residual",
                   "fast_path": f"mixed {args.inner} inner BiCGSTAB, FP64 reliable updates (delta "
                                f"{args.delta}); learned preconditioner: not applied",
                   "plain_path": "FP64 even-odd BiCGSTAB"},
        "tts_fast_s": round(t_fast, 4),
        "tts_plain_s": round(t_plain_med, 4),
        "iters_fast": fast_res[-1].iters,
        "reliable_updates_fast": fast_res[-1].reliable_updates,
        "iters_plain": plain_res.iters,
        "true_res_fast": fast_res[-1].true_res,
        "true_res_plain": plain_res.true_res,
        "roofline": {"kernel": f"hop {args.inner} clover xpay", "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_GBS, 3), "traffic": traffic,
                     "ms_per_launch": round(ms_hop, 5), "alg_bytes_per_launch": alg},
        "hop_fp64_clover_xpay": {"ms_per_launch": round(ms64, 5),
                                 "achieved_GBps": round(alg64 / (ms64 * 1e-3) / 1e9, 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
