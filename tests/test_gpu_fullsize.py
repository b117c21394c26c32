"""Checks at the full 32^3 x 64 lattice of the headline workload.

The small-lattice tests pin every kernel against the oracle; these repeat the
checks at the production shape, where launch grids, the XCD block remap and
the tiled wave order of the network kernel take their real sizes:
* the FP64 clover hop against the oracle's (OpenMP, FP64) to 1e-12;
* the matrix-core network (tiled wave order) against the vector-FMA network,
  itself pinned against the oracle on small lattices;
* a mixed-precision FP16 solve checked by the oracle's FP64 true residual.
The oracle takes seconds per call at this size (gauge generation, one hop, one
residual).
"""
import numpy as np
import pytest

from gpu_util import dev, host, log, rel

pytestmark = pytest.mark.gpu

L = (32, 32, 32, 64)
KAPPA, CSW, TB = 0.1248, 1.0, -1


@pytest.fixture(scope="module")
def full(orc):
    from qcu_amd.lattice import Lattice
    U = orc.gauge_random(L, 0.2, 900)
    lat = Lattice(L, KAPPA, CSW, TB)
    lat.load_gauge(dev(U))
    log("full-size lattice ready")
    return U, lat


def test_full_size_clover_hop_matches_oracle(orc, full):
    from qcu_amd.lattice import FP64
    U, lat = full
    psi = orc.gauss((orc.vol(L) // 2, 4, 3), 901)
    out = host(lat.dslash(dev(psi), parity=0, clover=True, prec=FP64))
    ref = orc.clover_dslash(psi, U, L, 0, KAPPA * CSW, TB)
    e = rel(out, ref)
    log(f"32^3x64 FP64 clover hop vs oracle: {e:.2e}")
    assert e < 1e-12


def test_full_size_network_kernels_agree(full, monkeypatch):
    """the matrix-core layers (tiled order at this shape) against the vector layers"""
    from qcu_amd.lattice import LearnedPreconditioner, gaussian
    _, lat = full
    chan = (1, 8, 8, 8, 1)
    net = LearnedPreconditioner(lat, chan, seed=902, init_scale=0.5)
    W = net.get_weights()
    x = gaussian((lat.vh, 4, 3), 903)
    y_mfma = host(net.apply_odd(x))
    net.close()
    monkeypatch.setenv("QCU_NET_KERNEL", "valu")
    ref = LearnedPreconditioner(lat, chan, seed=902, init_scale=0.5)
    ref.set_weights(W)
    y_valu = host(ref.apply_odd(x))
    ref.close()
    e = rel(y_mfma, y_valu)
    log(f"32^3x64 network, matrix-core vs vector kernels: {e:.2e}, |y| = {np.linalg.norm(y_valu):.3e}")
    # FP32 products, sums over 9 x 8 terms per layer in two different orders
    assert np.isfinite(y_mfma).all() and e < 1e-5


def test_full_size_mixed_solve_true_residual(orc, full):
    from qcu_amd.lattice import FP16
    U, lat = full
    b = orc.gauss((2, orc.vol(L) // 2, 4, 3), 904)
    x, res = lat.solve(dev(b), tol=1e-10, prec=FP16, reliable_delta=0.01)
    tr = orc.true_residual(host(x), b, U, L, KAPPA, CSW, TB)
    log(f"32^3x64 FP16 mixed solve: {res.iters} iters, {res.reliable_updates} reliable updates, "
        f"oracle true residual {tr:.2e}")
    assert res.converged and tr < 1e-10
