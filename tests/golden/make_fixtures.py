"""Writes the golden fixtures of the hot path from the FP64 host oracle.

    python tests/golden/make_fixtures.py      # rewrites tests/golden/*.json

The reference stores no input/output vectors for this path (its tests compare
against QUDA at run time), so these fixtures are outputs of the oracle, which
is pinned separately (closed forms, covariance, dense solves, an independent
numpy restatement).  They freeze, for seeded inputs, what every later build
must reproduce: norms, a weighted checksum, and the full 12 complex values at
a few sites.  Layouts are the external PyQuda ones.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle  # noqa: E402

SITES = [0, 1, 777, 2047]  # half-lattice site indices sampled in each parity


def checksum(a):
    """sum_k w_k (re_k + 2 im_k) with fixed pseudo-random weights; catches permutations"""
    a = np.ascontiguousarray(a).reshape(-1)
    w = ((np.arange(a.size, dtype=np.int64) * 2654435761) % 1000003) / 1000003.0
    return float(np.sum(w * (a.real + 2.0 * a.imag)))


def sites(f, n):
    f = f.reshape(n, -1, 12)
    return {f"{p}:{i}": [[float(z.real), float(z.imag)] for z in f[p, i]] for p in range(n) for i in SITES
            if i < f.shape[1]}


def c0():
    """8^4, eps 0.3, kappa 0.125, antiperiodic t: Wilson and clover hops"""
    L, eps, gseed, sseed = (8, 8, 8, 8), 0.3, 2024, 7
    U = oracle.gauge_random(L, eps, gseed)
    vh = oracle.vol(L) // 2
    psi = oracle.gauss((2, vh, 4, 3), sseed)
    out = {"L": L, "eps": eps, "gauge_seed": gseed, "spinor_seed": sseed, "kappa": 0.125, "t_boundary": -1,
           "gauge_checksum": checksum(U), "gauge_norm2": float(np.vdot(U, U).real),
           "spinor_checksum": checksum(psi)}
    hop = np.stack([oracle.hop(psi[1 - p], U, L, p, -1) for p in (0, 1)])
    out["hop"] = {"norm2": [float(np.vdot(h, h).real) for h in hop], "checksum": checksum(hop),
                  "sites": sites(hop, 2)}
    cl = np.stack([oracle.clover_dslash(psi[1 - p], U, L, p, 0.125, -1) for p in (0, 1)])
    out["clover_hop"] = {"coeff": 0.125, "norm2": [float(np.vdot(h, h).real) for h in cl],
                         "checksum": checksum(cl), "sites": sites(cl, 2)}
    return out


def solve4():
    """4^4 clover solve, kappa 0.124, csw 1, point source at the origin"""
    L, eps, gseed = (4, 4, 4, 4), 0.3, 99
    U = oracle.gauge_random(L, eps, gseed)
    vh = oracle.vol(L) // 2
    b = np.zeros((2, vh, 4, 3), complex)
    b[0, 0, 0, 0] = 1.0
    x, info = oracle.bicgstab(b, U, L, 0.124, 1.0, -1, tol=1e-12, max_iter=1000, mode=1)
    return {"L": L, "eps": eps, "gauge_seed": gseed, "kappa": 0.124, "csw": 1.0, "t_boundary": -1,
            "source": "point (0,0,0,0) spin 0 colour 0", "iters": info["iters"],
            "true_res": info["true_res"], "x_norm2": float(np.vdot(x, x).real), "x_checksum": checksum(x),
            "x_sites": sites(x, 2)}


def main():
    oracle.build()
    for name, fn in (("c0_hop.json", c0), ("solve_4x4x4x4_clover.json", solve4)):
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(fn(), f, indent=1)
        print("wrote", name)


if __name__ == "__main__":
    main()
