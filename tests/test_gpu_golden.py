"""The HIP path against the committed golden fixtures (tests/golden/*.json):
the FP64 Wilson and clover hops at 8^4 to 1e-12 and the 4^4 clover solve to
its stated solution."""
import json
import os

import numpy as np
import pytest

from gpu_util import dev, host, log

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


def worst(field, n, want):
    f = field.reshape(n, -1, 12)
    e = 0.0
    for key, v in want.items():
        p, i = map(int, key.split(":"))
        val = np.array([complex(a, b) for a, b in v])
        e = max(e, float(np.linalg.norm(f[p, i] - val) / np.linalg.norm(val)))
    return e


def test_c0_hops_match_fixture(orc):
    from qcu_amd.lattice import FP64, Lattice
    g = load("c0_hop.json")
    L = tuple(g["L"])
    U = orc.gauge_random(L, g["eps"], g["gauge_seed"])
    psi = orc.gauss((2, orc.vol(L) // 2, 4, 3), g["spinor_seed"])
    lat = Lattice(L, g["kappa"], 1.0, g["t_boundary"])
    lat.load_gauge(dev(U))
    hop = np.stack([host(lat.dslash(dev(psi[1 - p]), p, clover=False, prec=FP64)) for p in (0, 1)])
    cl = np.stack([host(lat.dslash(dev(psi[1 - p]), p, clover=True, prec=FP64)) for p in (0, 1)])
    e1, e2 = worst(hop, 2, g["hop"]["sites"]), worst(cl, 2, g["clover_hop"]["sites"])
    n1 = max(abs(np.vdot(hop[p], hop[p]).real / g["hop"]["norm2"][p] - 1) for p in (0, 1))
    log(f"c0 fixture: hop sites {e1:.1e}, norms {n1:.1e}; clover hop sites {e2:.1e}")
    assert e1 < 1e-12 and n1 < 1e-12 and e2 < 1e-12


def test_small_clover_solve_matches_fixture(orc):
    from qcu_amd.lattice import FP64, Lattice, point_source
    g = load("solve_4x4x4x4_clover.json")
    L = tuple(g["L"])
    U = orc.gauge_random(L, g["eps"], g["gauge_seed"])
    lat = Lattice(L, g["kappa"], g["csw"], g["t_boundary"])
    lat.load_gauge(dev(U))
    x, r = lat.solve(point_source(L), tol=1e-12, prec=FP64)
    e = worst(host(x), 2, g["x_sites"])
    log(f"4^4 clover solve: {r.iters} iters (oracle {g['iters']}), solution sites {e:.1e}")
    assert r.converged and e < 1e-10


def test_weak_field_configuration(orc):
    """the reference's weak-field ILDG configuration (4^3x8): FP64 Wilson and
    clover hops against the oracle, and a clover solve to 1e-10"""
    from qcu_amd.gauge_io import read_ildg
    from qcu_amd.lattice import FP16, FP64, Lattice
    U, L = read_ildg(os.path.join(HERE, "golden", "weak_field.lime"))
    L = tuple(L)
    U = np.ascontiguousarray(U.reshape(4, 2, -1, 3, 3))
    kappa, csw = 0.124, 1.0
    lat = Lattice(L, kappa, csw, -1)
    lat.load_gauge(dev(U))
    psi = orc.gauss((orc.vol(L) // 2, 4, 3), 3)
    for p in (0, 1):
        h = host(lat.dslash(dev(psi), p, clover=False, prec=FP64))
        c = host(lat.dslash(dev(psi), p, clover=True, prec=FP64))
        eh = np.linalg.norm(h - orc.hop(psi, U, L, p, -1)) / np.linalg.norm(h)
        ec = np.linalg.norm(c - orc.clover_dslash(psi, U, L, p, kappa * csw, -1)) / np.linalg.norm(c)
        assert eh < 1e-12 and ec < 1e-12, (p, eh, ec)
    b = dev(orc.gauss((2, orc.vol(L) // 2, 4, 3), 4))
    for prec in (FP64, FP16):
        x, r = lat.solve(b, tol=1e-10, prec=prec)
        tr = orc.true_residual(host(x), host(b), U, L, kappa, csw, -1)
        log(f"weak field, prec {prec}: {r.iters} iters, oracle true residual {tr:.1e}")
        assert r.converged and tr <= 1e-10
