"""The oracle against the committed golden fixtures (tests/golden/*.json).

The fixtures freeze seeded inputs and their outputs (norms, a weighted
checksum, full site values); the oracle must keep reproducing them.  The GPU
path is held to the same fixtures in test_gpu_golden.py.
"""
import json
import os

import numpy as np

from golden.make_fixtures import SITES, checksum  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


def site_array(d):
    return {k: np.array([complex(a, b) for a, b in v]) for k, v in d.items()}


def check_sites(field, n, want, tol):
    f = field.reshape(n, -1, 12)
    for key, val in site_array(want).items():
        p, i = map(int, key.split(":"))
        err = np.linalg.norm(f[p, i] - val) / np.linalg.norm(val)
        assert err < tol, (key, err)


def test_c0_inputs_and_hops(orc):
    g = load("c0_hop.json")
    L = tuple(g["L"])
    U = orc.gauge_random(L, g["eps"], g["gauge_seed"])
    psi = orc.gauss((2, orc.vol(L) // 2, 4, 3), g["spinor_seed"])
    assert abs(checksum(U) - g["gauge_checksum"]) < 1e-9 * abs(g["gauge_checksum"])
    assert abs(checksum(psi) - g["spinor_checksum"]) < 1e-9 * abs(g["spinor_checksum"])
    hop = np.stack([orc.hop(psi[1 - p], U, L, p, g["t_boundary"]) for p in (0, 1)])
    for p in (0, 1):
        assert abs(np.vdot(hop[p], hop[p]).real - g["hop"]["norm2"][p]) < 1e-12 * g["hop"]["norm2"][p]
    check_sites(hop, 2, g["hop"]["sites"], 1e-13)
    cl = np.stack([orc.clover_dslash(psi[1 - p], U, L, p, g["clover_hop"]["coeff"], g["t_boundary"])
                   for p in (0, 1)])
    check_sites(cl, 2, g["clover_hop"]["sites"], 1e-12)
    assert abs(checksum(cl) - g["clover_hop"]["checksum"]) < 1e-9 * abs(g["clover_hop"]["checksum"])


def test_small_clover_solve(orc):
    g = load("solve_4x4x4x4_clover.json")
    L = tuple(g["L"])
    U = orc.gauge_random(L, g["eps"], g["gauge_seed"])
    b = np.zeros((2, orc.vol(L) // 2, 4, 3), complex)
    b[0, 0, 0, 0] = 1.0
    x, info = orc.bicgstab(b, U, L, g["kappa"], g["csw"], g["t_boundary"], tol=1e-12, max_iter=1000, mode=1)
    assert info["iters"] == g["iters"]
    assert info["true_res"] < 1e-12
    check_sites(x, 2, g["x_sites"], 1e-10)
