"""Benchmarks of the smaller workloads (time-to-solution at 1e-10 true
residual) and the kernel measurements behind the rocprofv3 summaries.

    python -m qcu_amd.configbench [c1] [c2] [c3a] [c3b] [c4] [legacy]
    python -m qcu_amd.configbench hop [--prec P --clover 0/1 --xpay 0/1 --iters N]
    python -m qcu_amd.configbench net-layer [--layer K --iters N]
    python -m qcu_amd.configbench net-apply [--iters N]
    python -m qcu_amd.configbench net-stage [--layer K --iters N]
    python -m qcu_amd.configbench net-backward [--iters N]
    python -m qcu_amd.configbench copy [--mib M --iters N]

c1  plain clover BiCGSTAB FP64, 16^3x32, eps 0.3, csw 1.0, kappa 0.124, point
    source (origin, spin 0, colour 0): iterations, time, FP64 hop GB/s.
c2  24^3x48 clover, eps 0.25, csw 1.0, kappa 0.1245, Gaussian source: FP64,
    and FP32 / FP16 inner solves at reliable-update drops 0.1, 0.01, 0.001,
    with the error of each solution against the FP64 one.
c3a 12 spin-colour point sources on the c2 lattice in one batch (FP32 inner),
    against 12 single solves; the 12-RHS hop against the 1-RHS hop; the same
    batch with an FP16 inner solver.
c3b 16 Wilson lattices (4 each of 8^4, 12^4, 16^3x32, 20^3x40; eps 0.3,
    kappa 0.124, Gaussian sources) in one batch against 16 single solves.
c4  32^3x64 clover, eps 0.2, csw 1.0, kappa 0.1248, Gaussian source: plain
    FP64 and mixed FP16 (reliable-update drop 0.01) solves, no network (the
    solver part of the headline benchmark, e.g. for profiling).
legacy  the reference's entry points through qcu_amd.pyqcu on the c1 lattice
    (16^3x32, eps 0.3): dslashQcu, dslashCloverQcu and ncclDslashQcu per call
    (first call and steady state: the first call of a gauge pointer converts
    the links and, for clover, builds the clover term), against the resident
    FP64 hop of the handle API; ncclBistabCgQcu (kappa 0.125, periodic, stop
    on Re<r,r> < 1e-9) with its |b - M x|^2.

Kernel measurements, all on the 32^3x64 clover lattice of the headline
benchmark (eps 0.2, csw 1.0, kappa 0.1248) unless --L is given:
hop          the resident hop kernel (qcuBenchHop): HIP-event ms per launch and
             algorithmic GB/s.
net-layer    one forward layer of the 1-8-8-8-1 network (qcuNetBenchLayer) on
             the activations of a forward pass over Gaussian data.
net-apply    the network applied as the preconditioner (odd field in, odd part
             of the output kept; qcuNetApplyOdd), median wall ms.
net-stage    one layer's kernels as the preconditioner runs them
             (qcuNetBenchPrecondLayer), ms per launch in HIP events.
net-backward full forward + backward passes (qcuNetBackward), median wall ms.
copy         a device-to-device copy and a fill of M MiB: the plain streams of
             known size that scale the FETCH_SIZE / WRITE_SIZE counters.

Every solve must reach |b - M x| / |b| <= 1e-10 (FP64); one JSON line per
workload.  Times are medians of 3 (2 for batches) after a warm-up solve.
"""
import json
import sys
import time

import torch

from .lattice import FP16, FP32, FP64, Batch, Lattice, gaussian, point_source, random_gauge

# algorithmic bytes per output site of the hop, the clover inverse and the xpay operand
HOP = {0: 8 * 192 + 8 * 96 + 192, 1: 8 * 96 + 8 * 48 + 96, 2: 8 * 52 + 8 * 24 + 52}
CL = {0: 576, 1: 288, 2: 148}
XP = {0: 192, 1: 96, 2: 52}
TOL = 1e-10


def timed(fn, n=3):
    fn()
    ts, r = [], None
    for _ in range(n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[n // 2], r


def check(res):
    res = res if isinstance(res, list) else [res]
    for r in res:
        if not (r.converged and r.true_res <= TOL):
            raise RuntimeError(f"solve missed {TOL}: {r}")


def c1():
    L = (16, 16, 16, 32)
    lat = Lattice(L, 0.124, 1.0, -1)
    lat.load_gauge(random_gauge(L, 0.3, 11))
    lat.tune(FP64)
    b = point_source(L)
    t, (x, r) = timed(lambda: lat.solve(b, tol=TOL, prec=FP64))
    check(r)
    ms = lat.bench_hop(FP64, True, True, 100)
    return {"iters": r.iters, "tts_ms": t * 1e3, "true_res": r.true_res, "hop_fp64_clover_xpay_ms": ms,
            "GBps": (HOP[0] + CL[0] + XP[0]) * lat.vh / ms / 1e6, "GFLOPs": 1320 * lat.vh / ms / 1e6}


def c2():
    L = (24, 24, 24, 48)
    lat = Lattice(L, 0.1245, 1.0, -1)
    lat.load_gauge(random_gauge(L, 0.25, 12))
    for p in (FP64, FP32, FP16):
        lat.tune(p)
    b = gaussian((2, 48, 24, 24, 12, 4, 3), 13)
    t64, (x64, r64) = timed(lambda: lat.solve(b, tol=TOL, prec=FP64))
    check(r64)
    d = {"fp64": {"iters": r64.iters, "tts_ms": t64 * 1e3, "true_res": r64.true_res}}
    for name, p in (("fp32", FP32), ("fp16", FP16)):
        for delta in (0.1, 0.01, 0.001):
            t, (x, r) = timed(lambda: lat.solve(b, tol=TOL, prec=p, reliable_delta=delta))
            check(r)
            err = float(torch.linalg.norm(x - x64) / torch.linalg.norm(x64))
            d[f"{name}_d{delta}"] = {"iters": r.iters, "reliable": r.reliable_updates, "tts_ms": t * 1e3,
                                     "true_res": r.true_res, "rel_err_vs_fp64_solution": err}
    for p in (FP64, FP32, FP16):
        ms = lat.bench_hop(p, True, True, 100)
        d[f"hop_p{p}_ms"] = ms
        d[f"hop_p{p}_GBps"] = (HOP[p] + CL[p] + XP[p]) * lat.vh / ms / 1e6
    return d


def c3a():
    L = (24, 24, 24, 48)
    lat = Lattice(L, 0.1245, 1.0, -1)
    lat.load_gauge(random_gauge(L, 0.25, 12))
    lat.tune(FP32)
    bs = [point_source(L, s, c) for s in range(4) for c in range(3)]
    batch = Batch([lat], [12])
    batch.tune(FP32)
    tb, (xs, rs) = timed(lambda: batch.solve(bs, tol=TOL, prec=FP32, reliable_delta=0.01), 2)
    check(rs)
    ts, res = timed(lambda: [lat.solve(b, tol=TOL, prec=FP32, reliable_delta=0.01) for b in bs], 2)
    err = max(float(torch.linalg.norm(x - y[0]) / torch.linalg.norm(y[0])) for x, y in zip(xs, res))
    ms1 = lat.bench_hop(FP32, True, True, 50)
    ms12 = batch.bench_hop(FP32, True, 50)
    # the same batch with an FP16 inner solver
    lat.tune(FP16)
    batch.tune(FP16)
    th, (xh, rh) = timed(lambda: batch.solve(bs, tol=TOL, prec=FP16, reliable_delta=0.01), 2)
    check(rh)
    err16 = max(float(torch.linalg.norm(x - y[0]) / torch.linalg.norm(y[0])) for x, y in zip(xh, res))
    ms12h = batch.bench_hop(FP16, True, 50)
    return {"batch_ms": tb * 1e3, "sequential_ms": ts * 1e3, "iters": [r.iters for r in rs],
            "max_true_res": max(r.true_res for r in rs), "max_rel_diff_vs_single": err,
            "hop_fp32_1rhs_ms": ms1, "hop_fp32_12rhs_ms": ms12, "hop_12rhs_per_rhs_ms": ms12 / 12,
            "fp16": {"batch_ms": th * 1e3, "iters": [r.iters for r in rh], "max_true_res": max(r.true_res for r in rh),
                     "max_rel_diff_vs_fp32_single": err16, "hop_fp16_12rhs_ms": ms12h}}


def c3b():
    shapes = [(8, 8, 8, 8)] * 4 + [(12, 12, 12, 12)] * 4 + [(16, 16, 16, 32)] * 4 + [(20, 20, 20, 40)] * 4
    lats, bs = [], []
    for k, L in enumerate(shapes):
        lat = Lattice(L, 0.124, 0.0, -1)
        lat.load_gauge(random_gauge(L, 0.3, 100 + k))
        lats.append(lat)
        bs.append(gaussian((2, L[3], L[2], L[1], L[0] // 2, 4, 3), 200 + k))
    batch = Batch(lats, [1] * 16)
    for p in (FP64, FP32):
        batch.tune(p)
        for lat in lats:
            lat.tune(p)
    d = {}
    for name, p, delta in (("fp64", FP64, 0.1), ("fp32", FP32, 0.01)):
        tb, (xs, rs) = timed(lambda: batch.solve(bs, tol=TOL, prec=p, reliable_delta=delta), 2)
        check(rs)
        ts, _ = timed(lambda: [lat.solve(b, tol=TOL, prec=p, reliable_delta=delta) for lat, b in zip(lats, bs)], 2)
        d[name] = {"batch_ms": tb * 1e3, "sequential_ms": ts * 1e3, "iters": [r.iters for r in rs],
                   "max_true_res": max(r.true_res for r in rs)}
    return d


def c4():
    L = (32, 32, 32, 64)
    lat = Lattice(L, 0.1248, 1.0, -1)
    lat.load_gauge(random_gauge(L, 0.2, 20241))
    b = gaussian((2, 64, 32, 32, 16, 4, 3), 777)
    for p in (FP64, FP16):
        lat.tune(p)
    d = {}
    for name, p, delta in (("fp64", FP64, 0.1), ("fp16", FP16, 0.01)):
        t, (x, r) = timed(lambda: lat.solve(b, tol=TOL, prec=p, reliable_delta=delta))
        check(r)
        d[name] = {"iters": r.iters, "reliable": r.reliable_updates, "tts_ms": t * 1e3, "true_res": r.true_res}
    return d


def legacy():
    from . import pyqcu
    from .lattice import fermion_shape
    L = (16, 16, 16, 32)
    U = random_gauge(L, 0.3, 12)
    param = pyqcu.QcuParam()
    param.lattice_size = list(L)
    grid = pyqcu.QcuParam()
    grid.lattice_size = [1, 1, 1, 1]
    psi = gaussian(fermion_shape(L, full=False), 13)
    out = torch.empty_like(psi)

    def per_call(fn, n=7):
        ts = []
        for _ in range(n):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return ts[0] * 1e3, sorted(ts[1:])[(n - 1) // 2] * 1e3

    d = {}
    for name, fn in (("dslashQcu", lambda: pyqcu.dslashQcu(out, psi, U, param, 0)),
                     ("dslashCloverQcu", lambda: pyqcu.dslashCloverQcu(out, psi, U, param, 0)),
                     ("ncclDslashQcu", lambda: pyqcu.ncclDslashQcu(out, psi, U, param, 1, grid))):
        first, steady = per_call(fn)
        d[name] = {"first_ms": first, "steady_ms": steady}
    lat = Lattice(L, 0.125, 0.0, 1)
    lat.load_gauge(U)
    d["resident_fp64_hop_ms"] = lat.bench_hop(prec=FP64, clover=False, xpay=False, iters=50)
    b = gaussian(fermion_shape(L), 14)
    x = torch.zeros_like(b)
    t, _ = timed(lambda: pyqcu.ncclBistabCgQcu(x, b, U, param, grid))
    # the handle API's operator at the legacy kappa and boundary gives |b - M x|^2
    r = b - lat.mat(x)
    d["ncclBistabCgQcu"] = {"tts_ms": t * 1e3, "r2": float(torch.vdot(r.flatten(), r.flatten()).real)}
    lat.close()
    return d


def _headline_lattice(args):
    L = tuple(args.L)
    lat = Lattice(L, 0.1248, 1.0 if args.clover else 0.0, -1)
    U = random_gauge(L, 0.2, 1)
    lat.load_gauge(U)
    del U
    torch.cuda.empty_cache()
    return lat


def _median_ms(fn, n):
    ts = []
    for _ in range(n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    return sorted(ts)[len(ts) // 2]


def hop(args):
    lat = _headline_lattice(args)
    if args.shape:
        lat.set_launch_shape(args.prec, *args.shape)
    elif args.tune:
        lat.tune(args.prec)
    ms = lat.bench_hop(prec=args.prec, clover=bool(args.clover), xpay=bool(args.xpay), iters=args.iters)
    alg = (HOP[args.prec] + (CL[args.prec] if args.clover else 0) + (XP[args.prec] if args.xpay else 0)) * lat.vh
    return {"L": lat.L, "prec": args.prec, "clover": args.clover, "xpay": args.xpay,
            "shape": lat.launch_shape(args.prec), "ms_per_launch": ms, "alg_bytes": alg,
            "alg_GBps": alg / ms / 1e6}


def _network(lat, seed=3, init_scale=0.3):
    from .lattice import LearnedPreconditioner
    return LearnedPreconditioner(lat, (1, 8, 8, 8, 1), seed=seed, init_scale=init_scale)


def net_layer(args):
    from .lattice import fermion_shape, gaussian
    lat = _headline_lattice(args)
    net = _network(lat, init_scale=0.05)
    # one forward pass on Gaussian data fills every layer's input with
    # realistic activations (all-zero operands draw less power and run the
    # matrix cores at a higher clock than real data)
    net.forward(gaussian(fermion_shape(lat.L), 7))
    ms, fl = net.bench_layer(args.layer, iters=args.iters)
    return {"L": lat.L, "net_layer": args.layer, "ms_per_launch": ms, "alg_flops": fl, "TFLOPs": fl / ms / 1e9}


def net_apply(args):
    from .lattice import gaussian
    lat = _headline_lattice(args)
    net = _network(lat)
    x = gaussian((lat.vh, 4, 3), 7)
    for _ in range(3):
        net.apply_odd(x)
    return {"L": lat.L, "net_apply_ms_median": _median_ms(lambda: net.apply_odd(x), args.iters),
            "net_apply_ms_events": net.bench_apply(iters=args.iters)}


def net_stage(args):
    lat = _headline_lattice(args)
    net = _network(lat)
    return {"L": lat.L, "layer": args.layer, "ms_per_launch": net.bench_precond_layer(args.layer, iters=args.iters)}


def net_backward(args):
    from .lattice import fermion_shape, gaussian
    lat = _headline_lattice(args)
    net = _network(lat)
    x = gaussian((1,) + fermion_shape(lat.L), 5)
    gout = gaussian((1,) + fermion_shape(lat.L), 6)
    net.backward(x, gout)
    return {"L": lat.L, "net_backward_ms_median": _median_ms(lambda: net.backward(x, gout), args.iters)}


def copy(args):
    n = args.mib * (1 << 20)
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    b.fill_(1)
    torch.cuda.synchronize()

    def per_launch(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.iters):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.iters

    ms = per_launch(lambda: b.copy_(a))
    ms_fill = per_launch(lambda: b.fill_(1))
    return {"copy_bytes_read": n, "copy_bytes_written": n, "ms": ms, "GBps_rw": 2 * n / ms / 1e6,
            "fill_ms": ms_fill, "fill_GBps": n / ms_fill / 1e6}


def main():
    import argparse
    fns = {"c1": c1, "c2": c2, "c3a": c3a, "c3b": c3b, "c4": c4, "legacy": legacy}
    kernels = {"hop": hop, "net-layer": net_layer, "net-apply": net_apply, "net-stage": net_stage, "net-backward": net_backward,
               "copy": copy}
    ap = argparse.ArgumentParser(description="workload and kernel measurements")
    ap.add_argument("which", nargs="*", default=["c1", "c2", "c3a", "c3b"],
                    choices=sorted(fns) + sorted(kernels))
    ap.add_argument("--L", type=int, nargs=4, default=[32, 32, 32, 64])
    ap.add_argument("--prec", type=int, default=2)
    ap.add_argument("--clover", type=int, default=1)
    ap.add_argument("--xpay", type=int, default=1)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--layer", type=int, default=1)
    ap.add_argument("--tune", type=int, default=1, help="hop: tune the launch shape first (0: default shape)")
    ap.add_argument("--shape", type=int, nargs=3, metavar=("BLOCK", "REMAP", "GROUP"),
                    help="hop: run this launch shape (as the tuner would pick it) instead of tuning")
    ap.add_argument("--mib", type=int, default=1024)
    args = ap.parse_args()
    for w in args.which:
        res = kernels[w](args) if w in kernels else fns[w]()
        print(json.dumps({"workload": w, **res}), flush=True)


if __name__ == "__main__":
    main()
