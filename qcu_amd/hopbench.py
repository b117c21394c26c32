"""Hop-kernel microbenchmark (profiling driver).

    python -m qcu_amd.hopbench --prec 2 --clover 1 --xpay 1 --iters 50
    python -m qcu_amd.hopbench --copy-mib 1024 --iters 50   # plain copy of known size
    python -m qcu_amd.hopbench --net-layer 1 --iters 5       # learned-network layer 8 -> 8
    python -m qcu_amd.hopbench --net-backward 5              # full forward + backward passes
    python -m qcu_amd.hopbench --net-apply 20                # preconditioner applications

Runs the resident hop kernel on a 32^3x64 clover lattice (eps 0.2, csw 1.0,
kappa 0.1248) and prints the HIP-event time per launch and the algorithmic
GB/s.  The copy mode launches a device-to-device copy of a known byte count,
used to calibrate the FETCH_SIZE / WRITE_SIZE counters.  The net-layer mode
times one forward layer (both parities) of the 1-8-8-8-1 channel network.
The net-apply mode times the network applied as the preconditioner (an odd
field in, the odd part of the output kept), as in the learned solve.
"""
import argparse
import json

HOP = {0: 8 * 192 + 8 * 96 + 192, 1: 8 * 96 + 8 * 48 + 96, 2: 8 * 52 + 8 * 24 + 52}
CLOVER = {0: 576, 1: 288, 2: 148}
XPAY = {0: 192, 1: 96, 2: 52}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, nargs=4, default=[32, 32, 32, 64])
    ap.add_argument("--prec", type=int, default=2)
    ap.add_argument("--clover", type=int, default=1)
    ap.add_argument("--xpay", type=int, default=1)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--remap", type=int, default=-1)
    ap.add_argument("--copy-mib", type=int, default=0)
    ap.add_argument("--net-layer", type=int, default=-1)
    ap.add_argument("--net-backward", type=int, default=0, help="time N full backward passes of the network")
    ap.add_argument("--net-apply", type=int, default=0, help="time N preconditioner applications of the network")
    args = ap.parse_args()

    import torch
    if args.copy_mib:
        n = args.copy_mib * (1 << 20)
        a = torch.empty(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.iters):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.iters
        # a write-only stream of the same size (one untimed call loads the kernel)
        b.fill_(1)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(args.iters):
            b.fill_(1)
        e1.record()
        torch.cuda.synchronize()
        ms_fill = e0.elapsed_time(e1) / args.iters
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "ms": ms,
                          "GBps_rw": 2 * n / ms / 1e6, "fill_ms": ms_fill, "fill_GBps": n / ms_fill / 1e6}))
        return

    from .lattice import Lattice, random_gauge
    L = tuple(args.L)
    U = random_gauge(L, 0.2, 1)
    lat = Lattice(L, 0.1248, 1.0 if args.clover else 0.0, -1)
    lat.load_gauge(U)
    del U
    torch.cuda.empty_cache()
    if args.net_apply > 0:
        import time
        from .lattice import LearnedPreconditioner, gaussian
        net = LearnedPreconditioner(lat, (1, 8, 8, 8, 1), seed=3, init_scale=0.3)
        x = gaussian((lat.vh, 4, 3), 7)
        for _ in range(3):
            net.apply_odd(x)
        ts = []
        for _ in range(args.net_apply):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            net.apply_odd(x)
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        print(json.dumps({"L": L, "net_apply_ms_median": sorted(ts)[len(ts) // 2]}))
        return
    if args.net_backward > 0:
        import time
        from .lattice import LearnedPreconditioner, fermion_shape, gaussian
        net = LearnedPreconditioner(lat, (1, 8, 8, 8, 1), seed=3, init_scale=0.3)
        x = gaussian((1,) + fermion_shape(L), 5)
        gout = gaussian((1,) + fermion_shape(L), 6)
        net.backward(x, gout)
        ts = []
        for _ in range(args.net_backward):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            net.backward(x, gout)
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        print(json.dumps({"L": L, "net_backward_ms_median": sorted(ts)[len(ts) // 2]}))
        return
    if args.net_layer >= 0:
        from .lattice import LearnedPreconditioner, fermion_shape, gaussian
        net = LearnedPreconditioner(lat, (1, 8, 8, 8, 1), seed=3)
        # one forward pass on Gaussian data fills every layer's input with
        # realistic activations (all-zero operands draw less power and run the
        # matrix cores at a higher clock than real data)
        net.forward(gaussian(fermion_shape(L), 7))
        ms, fl = net.bench_layer(args.net_layer, iters=args.iters)
        print(json.dumps({"L": L, "net_layer": args.net_layer, "ms_per_launch": ms, "alg_flops": fl,
                          "TFLOPs": fl / ms / 1e9}))
        return
    ms = lat.bench_hop(prec=args.prec, clover=bool(args.clover), xpay=bool(args.xpay), iters=args.iters,
                       block=args.block, remap=args.remap)
    alg = (HOP[args.prec] + (CLOVER[args.prec] if args.clover else 0) +
           (XPAY[args.prec] if args.xpay else 0)) * lat.vh
    print(json.dumps({"L": L, "prec": args.prec, "clover": args.clover, "xpay": args.xpay,
                      "ms_per_launch": ms, "alg_bytes": alg, "alg_GBps": alg / ms / 1e6}))


if __name__ == "__main__":
    main()
