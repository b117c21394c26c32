"""Headline benchmark: time-to-solution at 1e-10 true residual on a 32^3x64
clover lattice (gauge eps 0.2, csw 1.0, kappa 0.1248, antiperiodic t) of the
learned-preconditioned mixed-precision solve (4-layer gauge-covariant network
trained here on the GPU, FP16-with-scale inner BiCGSTAB, FP64 reliable
updates) against the plain FP64 even-odd BiCGSTAB on the same GPU.

A step is one complete learned+mixed solve of one Gaussian source already
resident in HBM (external PyQuda layout).  The plain FP64 solve and the mixed
solve without the network are timed alongside (median of `steps` solves);
value = t_plain / t_learned_mixed.  Training time is reported, not timed in
the step.

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
PEAK_FP32_TFLOPS = 157.3
PEAK_FP16_TFLOPS = 2500.0
# algorithmic bytes per output site of one hop launch (8 neighbour spinors,
# 8 compressed links, 1 output) and of the packed clover inverse, per precision
HOP_BYTES = {0: 8 * 192 + 8 * 96 + 192, 1: 8 * 96 + 8 * 48 + 96, 2: 8 * 52 + 8 * 24 + 52}
CLOVER_BYTES = {0: 576, 1: 288, 2: 148}
# the xpay epilogue of the Schur operator reads one more spinor
XPAY_BYTES = {0: 192, 1: 96, 2: 52}
# compulsory bytes per output site of the clover + xpay hop, every field once:
# the neighbour spinor field S, the 8 link slabs 8 L, the clover inverse, the
# xpay operand S and the output S (S, L = 192, 96 / 96, 48 / 52, 24 bytes)
COMPULSORY_HOP_BYTES = {0: 192 + 8 * 96 + 576 + 192 + 192, 1: 96 + 8 * 48 + 288 + 96 + 96,
                        2: 52 + 8 * 24 + 148 + 52 + 52}
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "hop_fp16_clover_32x32x32x64_traffic.json")
NET_TRAFFIC = os.path.join(ROOT, "profiles", "net_layer_8x8_split_32x32x32x64_traffic.json")
SECOND_TRAFFIC = os.path.join(ROOT, "profiles", "net_second_32x32x32x64_traffic.json")
HOP64_TRAFFIC = os.path.join(ROOT, "profiles", "hop_fp64_clover_32x32x32x64_traffic.json")
HOP64W_TRAFFIC = os.path.join(ROOT, "profiles", "hop_fp64_wilson_32x32x32x64_traffic.json")


def traffic_of(path):
    """HBM bytes per launch recorded from the PMC runs of profiles/ (None if absent)"""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info():
    """the host cores the CPU baseline ran on"""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"os_cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "cpu_model": model}


def cpu_baseline(U_host, b_host, threads, checks):
    """The FP64 OpenMP host oracle solves the same 32^3x64 clover system to the
    same 1e-10 true residual (measured, not extrapolated), and recomputes the
    true residual of each timed GPU solution in `checks` {name: x}."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.build()
    if threads:
        oracle.set_threads(threads)
    nthreads = oracle.threads()
    t0 = time.perf_counter()
    _, info = oracle.bicgstab(b_host, U_host, L, KAPPA, CSW, TB, tol=0.5 * TOL)
    wall = time.perf_counter() - t0
    # the host Wilson hop alone (one parity), median of 3
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.hop(b_host[1], U_host, L, 0, TB)
        ts.append(time.perf_counter() - t0)
    hop_s = sorted(ts)[1]
    vh = b_host.shape[1]
    per_iter = info["iter_s"] / max(info["iters"], 1)
    true_res = {name: oracle.true_residual(x, b_host, U_host, L, KAPPA, CSW, TB) for name, x in checks.items()}
    return {"value": round(info["setup_s"] + info["iter_s"], 3), "unit": "s (FP64 time-to-solution, measured)",
            "cores": nthreads, "kind": "port",
            "sample": f"one full FP64 even-odd BiCGSTAB solve of the same 32^3x64 clover system and source "
                      f"to true residual {info['true_res']:.2e} ({info['iters']} iterations, clover set-up "
                      f"{info['setup_s']:.1f} s, iterations {info['iter_s']:.1f} s; {wall:.1f} s wall with I/O)",
            "iters": info["iters"], "true_res": info["true_res"], "s_per_iteration": round(per_iter, 4),
            "omp_threads": nthreads, "host": host_info(),
            "hop_fp64": {"ms": round(hop_s * 1e3, 2), "GBps": round(HOP_BYTES[0] * vh / hop_s / 1e9, 2),
                         "GFLOPs": round(1320 * vh / hop_s / 1e9, 2)},
            "oracle_true_res_of_gpu_solutions": true_res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--inner", default="fp16", choices=["fp32", "fp16"])
    ap.add_argument("--delta", type=float, default=0.01, help="reliable-update residual drop")
    ap.add_argument("--train-sources", type=int, default=256)
    ap.add_argument("--train-epochs", type=int, default=2)
    ap.add_argument("--train-batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--solution-weight", type=float, default=0.0,
                    help="weight of the (source, solution) pair loss |K c - y|^2/|y|^2 in training "
                         "(0: the residual loss |Mhat K c - c|^2/|c|^2 only)")
    ap.add_argument("--pairs", action="store_true",
                    help="generate the (source, solution) pairs with the plain FP64 solver (counted in the "
                         "training cost) even when the loss does not use the solutions")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU halo exchange is out of scope)", file=sys.stderr)
        sys.exit(2)

    import torch
    from qcu_amd.lattice import (FP16, FP32, FP64, Lattice, LearnedPreconditioner, gaussian, random_gauge,
                                 schur_solutions)

    inner = FP16 if args.inner == "fp16" else FP32
    U = random_gauge(L, EPS, 20241)
    lat = Lattice(L, KAPPA, CSW, TB)
    lat.load_gauge(U)
    vh = lat.vh
    b = gaussian((2,) + (L[3], L[2], L[1], L[0] // 2, 4, 3), 777)
    torch.cuda.synchronize()
    log("gauge + clover resident; tuning hop launch shapes")
    for p in (FP64, inner):
        log(f"  prec {p}: {lat.tune(p)}")

    def timed(fn, n):
        ts, res, x = [], None, None
        for _ in range(n):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x, res = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
            assert res.converged and res.true_res <= TOL, f"solve did not reach {TOL}: {res}"
        return sorted(ts)[len(ts) // 2], res, x

    # plain FP64 BiCGSTAB and the mixed-precision solve without the network
    plain = lambda: lat.solve(b, tol=TOL, prec=FP64)
    mixed = lambda: lat.solve(b, tol=TOL, prec=inner, reliable_delta=args.delta)
    plain()
    mixed()
    t_plain, plain_res, x_plain = timed(plain, max(1, args.steps))
    log(f"plain FP64: {t_plain * 1e3:.1f} ms {plain_res}")
    t_mixed, mixed_res, _ = timed(mixed, max(1, args.steps))
    log(f"mixed {args.inner}: {t_mixed * 1e3:.1f} ms {mixed_res}")

    # the learned preconditioner, trained on the GPU on Gaussian odd-site sources
    net = LearnedPreconditioner(lat, (1, 8, 8, 8, 1), seed=3)
    srcs = [gaussian((vh, 4, 3), 10000 + k) for k in range(args.train_sources)]
    torch.cuda.synchronize()
    pair_s = 0.0
    sols = None
    if args.solution_weight > 0 or args.pairs:
        # (source, solution) pairs from the plain FP64 solver, y = Mhat^-1 c
        t0 = time.perf_counter()
        sols = schur_solutions(lat, srcs, tol=1e-8, prec=FP64)
        torch.cuda.synchronize()
        pair_s = time.perf_counter() - t0
    if args.solution_weight > 0:
        losses, train_s = net.train_pairs(srcs, sols, epochs=args.train_epochs, batch=args.train_batch,
                                          lr=args.lr, solution_weight=args.solution_weight)
    else:
        losses, train_s = net.train(srcs, epochs=args.train_epochs, batch=args.train_batch, lr=args.lr)
    del sols
    del srcs
    log(f"training: {train_s:.1f} s, epoch losses {losses.tolist()}")
    sol_pairs = args.solution_weight > 0 or args.pairs
    learned = lambda: lat.solve(b, tol=TOL, prec=inner, reliable_delta=args.delta, precond=net)
    for _ in range(args.warmup):
        learned()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lres = []
    for _ in range(args.steps):
        x_learned, r = learned()
        lres.append(r)
    torch.cuda.synchronize()
    t_learned = (time.perf_counter() - t0) / args.steps
    assert all(r.converged and r.true_res <= TOL for r in lres), "learned solve did not reach 1e-10"
    log(f"learned+mixed: {t_learned * 1e3:.1f} ms {lres[-1]}")

    # roofline of the dominant kernel of the timed solve: the network's second
    # layer as the preconditioner runs it (k_net_second<8, true>): its even
    # outputs from the one-channel input with combined matrices, its odd
    # outputs from the first layer's even sites and the input, and in its
    # epilogue the last layer's term fields for the third layer's odd sites;
    # split FP16 on the matrix cores, timed with HIP events on the stream it
    # runs on.  Its bound is memory: the compulsory bytes are the input's odd
    # sites, the first layer's 8 channels at even sites, the FP32 links once,
    # its 8 output channels (both parities) and the 9 term fields it writes;
    # the measured traffic (profiles/) shows the neighbour channels re-read.
    stages = [net.bench_precond_layer(k, iters=10) for k in range(4)]
    ms_second = stages[1]
    second_bytes = (96 + 8 * 96 + 8 * 48 + 2 * 8 * 96 + 9 * 96) * vh
    second_GBps = second_bytes / (ms_second * 1e-3) / 1e9
    second_traffic = traffic_of(SECOND_TRAFFIC)
    # executed v_mfma_f32_32x32x16_f16 (32768 flops each) per 32 sites of
    # both parities: even pass 4 K steps, odd pass 33, 6 products each per
    # M tile pair x 3 colours (18 per K step), epilogue 72 + 36
    second_mfma_flops = (4 * 18 + 33 * 18 + 72 + 36) * 32768 // 32 * vh
    # the plain 8 -> 8 layer (training's forward pass; every 8 -> 8 layer outside the preconditioner)
    ms_layer, fl_layer = net.bench_layer(1, iters=10)
    layer_bytes = 8 * 2 * vh * 96 * 2 + 8 * vh * 48
    layer_GBps = layer_bytes / (ms_layer * 1e-3) / 1e9
    net_traffic = traffic_of(NET_TRAFFIC)
    # executed matrix-core work: three FP16 products of the real-embedded
    # complex mixing, 9 terms x 8 x 8 channels x 16 spin pairs x 3 colours x
    # 4 real multiply-adds x 2 flops per site
    mfma_flops = 3 * 9 * 8 * 8 * 16 * 3 * 4 * 2 * 2 * vh
    ms_apply = net.bench_apply(iters=10)
    # the inner-precision clover hop with the xpay epilogue (HBM-bound)
    ms_hop = lat.bench_hop(prec=inner, clover=True, xpay=True, iters=100)
    alg = (HOP_BYTES[inner] + CLOVER_BYTES[inner] + XPAY_BYTES[inner]) * vh
    comp = COMPULSORY_HOP_BYTES[inner] * vh
    achieved = comp / (ms_hop * 1e-3) / 1e9
    traffic = traffic_of(PROFILE_TRAFFIC)
    traffic64 = traffic_of(HOP64_TRAFFIC)
    traffic64w = traffic_of(HOP64W_TRAFFIC)
    ms64 = lat.bench_hop(prec=FP64, clover=True, xpay=True, iters=50)
    alg64 = (HOP_BYTES[0] + CLOVER_BYTES[0] + XPAY_BYTES[0]) * vh
    ms64_plain = lat.bench_hop(prec=FP64, clover=False, xpay=False, iters=50)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline: the oracle's FP64 solve, and the oracle's true residual of the GPU solutions")
        cpu = cpu_baseline(U.cpu().numpy().reshape(4, 2, vh, 3, 3), b.cpu().numpy().reshape(2, vh, 4, 3),
                           args.cpu_threads,
                           {"plain": x_plain.cpu().numpy().reshape(2, vh, 4, 3),
                            "learned_mixed": x_learned.cpu().numpy().reshape(2, vh, 4, 3)})
        for name, tr in cpu["oracle_true_res_of_gpu_solutions"].items():
            assert tr <= TOL, f"the oracle's true residual of the {name} solution is {tr:.3e}"

    out = {
        "metric": METRIC,
        "value": round(t_plain / t_learned, 4),
        "unit": "x (plain FP64 time-to-solution / learned+mixed time-to-solution)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_learned * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (outer) / " + args.inner + " (inner) / f32 (network)",
        "data": "synthetic: U = exp(i 0.2 H) seeded, Gaussian source; network trained here on "
                f"{args.train_sources} Gaussian sources"
                + (" with their plain FP64 solutions" if sol_pairs else ""),
        "config": {"workload": "32^3x64 clover, eps 0.2, csw 1.0, kappa 0.1248, antiperiodic t, "
                               "tol 1e-10 true residual",
                   "fast_path": f"learned preconditioner (4 layers, 8 channels, FP32) + mixed "
                                f"{args.inner} inner BiCGSTAB, FP64 reliable updates (delta {args.delta})",
                   "plain_path": "FP64 even-odd BiCGSTAB"},
        "tts_learned_mixed_s": round(t_learned, 4),
        "tts_mixed_s": round(t_mixed, 4),
        "tts_plain_s": round(t_plain, 4),
        "speedup_mixed_only": round(t_plain / t_mixed, 3),
        "iters_learned_mixed": lres[-1].iters,
        "iters_mixed": mixed_res.iters,
        "iters_plain": plain_res.iters,
        "true_res": {"learned_mixed": lres[-1].true_res, "mixed": mixed_res.true_res,
                     "plain": plain_res.true_res},
        "training": {"seconds": round(train_s, 2), "sources": args.train_sources,
                     "loss": "residual |Mhat K c - c|^2/|c|^2" if args.solution_weight <= 0 else
                             f"{args.solution_weight} x pair |K c - y|^2/|y|^2 + {1 - args.solution_weight:g} x residual",
                     "pair_generation_seconds": round(pair_s, 2),
                     "pairs": "plain FP64 BiCGSTAB to 1e-8" if sol_pairs else None,
                     "epochs": args.train_epochs, "batch": args.train_batch,
                     "loss_first_epoch": float(losses[0]), "loss_last_epoch": float(losses[-1])},
        "roofline": {"kernel": "network layer 2 as the preconditioner runs it (k_net_second<8, true>)",
                     "bound": "hbm",
                     "compute": "v_mfma_f32_32x32x16_f16, FP32 operands split into FP16 hi + lo, three products "
                                "accumulated in FP32 (FP32-accurate)",
                     "bytes_count": "compulsory: the input's odd sites, layer 1's 8 channels at even sites, FP32 "
                                    "links once, 8 output channels (both parities), the 9 term fields written",
                     "achieved": round(second_GBps, 1), "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": round(second_GBps / PEAK_GBS, 3), "traffic": second_traffic,
                     "traffic_GBps": None if second_traffic is None else
                     round(second_traffic / (ms_second * 1e-3) / 1e9, 1),
                     "ms_per_launch": round(ms_second, 4), "alg_bytes_per_launch": second_bytes,
                     "fp16_mfma_TFLOPs": round(second_mfma_flops / (ms_second * 1e-3) / 1e12, 1),
                     "fp16_mfma_frac": round(second_mfma_flops / (ms_second * 1e-3) / 1e12 / PEAK_FP16_TFLOPS, 3)},
        "precond_layer_ms": {"layer1_first": round(stages[0], 4), "layer2_second": round(stages[1], 4),
                             "layer3_fused_even_sites": round(stages[2], 4), "layer4_gather": round(stages[3], 4)},
        "layer_8x8_plain": {"kernel": "network layer 8->8 channels (k_net_layer_split<32, 8, 2>)", "bound": "hbm",
                     "compute": "v_mfma_f32_32x32x16_f16, FP32 operands split into FP16 hi + lo, three products "
                                "accumulated in FP32 (FP32-accurate)",
                     "bytes_count": "compulsory: 8 input channels + 8 output channels (FP32, both parities) "
                                    "+ FP32 links once",
                     "achieved": round(layer_GBps, 1), "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": round(layer_GBps / PEAK_GBS, 3), "traffic": net_traffic,
                     "traffic_GBps": None if net_traffic is None else round(net_traffic / (ms_layer * 1e-3) / 1e9, 1),
                     "ms_per_launch": round(ms_layer, 4), "alg_bytes_per_launch": layer_bytes,
                     "fp32_equivalent_TFLOPs": round(fl_layer / (ms_layer * 1e-3) / 1e12, 1),
                     "fp16_mfma_TFLOPs": round(mfma_flops / (ms_layer * 1e-3) / 1e12, 1),
                     "fp16_mfma_frac": round(mfma_flops / (ms_layer * 1e-3) / 1e12 / PEAK_FP16_TFLOPS, 3)},
        "net_apply_ms": round(ms_apply, 4),
        "hop_roofline": {"kernel": f"hop {args.inner} clover xpay", "bound": "hbm",
                         "achieved": round(achieved, 1), "peak": PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / PEAK_GBS, 3), "traffic": traffic,
                         "hbm_GBps": None if traffic is None else round(traffic / (ms_hop * 1e-3) / 1e9, 1),
                         "hbm_frac": None if traffic is None else round(traffic / (ms_hop * 1e-3) / 1e9 / PEAK_GBS, 3),
                         "alg_GBps": round(alg / (ms_hop * 1e-3) / 1e9, 1),
                         "alg_frac": round(alg / (ms_hop * 1e-3) / 1e9 / PEAK_GBS, 3),
                         "note": "achieved counts each field once (compulsory bytes); alg counts each of the 8 "
                                 "neighbour spinors per site as the stencil reads it, so alg_frac can exceed 1 "
                                 "(neighbours shared through L2); hbm_frac is measured traffic / time / peak",
                         "ms_per_launch": round(ms_hop, 5), "compulsory_bytes_per_launch": comp,
                         "alg_bytes_per_launch": alg},
        "hop_fp64_wilson": {"ms_per_launch": round(ms64_plain, 5),
                            "achieved_GBps": round(HOP_BYTES[0] * vh / (ms64_plain * 1e-3) / 1e9, 1),
                            "GFLOPs": round(1320 * vh / (ms64_plain * 1e-3) / 1e9, 1),
                            "traffic": traffic64w,
                            "hbm_GBps": None if traffic64w is None else round(traffic64w / (ms64_plain * 1e-3) / 1e9, 1),
                            "hbm_frac": None if traffic64w is None else
                            round(traffic64w / (ms64_plain * 1e-3) / 1e9 / PEAK_GBS, 3)},
        "hop_fp64_clover_xpay": {"ms_per_launch": round(ms64, 5),
                                 "achieved_GBps": round(alg64 / (ms64 * 1e-3) / 1e9, 1),
                                 "compulsory_GBps": round(COMPULSORY_HOP_BYTES[0] * vh / (ms64 * 1e-3) / 1e9, 1),
                                 "traffic": traffic64,
                                 "hbm_GBps": None if traffic64 is None else round(traffic64 / (ms64 * 1e-3) / 1e9, 1),
                                 "hbm_frac": None if traffic64 is None else
                                 round(traffic64 / (ms64 * 1e-3) / 1e9 / PEAK_GBS, 3)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
