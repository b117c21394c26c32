"""Host-side checks of the propagator sources (qcu_amd.propagator), against the
checkerboard placement of python/PyQuda/pyquda/utils/source.py:7-48."""
import numpy as np
import torch

import indep
from qcu_amd import propagator as prop

L = (4, 6, 2, 8)


def test_point_source_position():
    for t_srce in ([0, 0, 0, 0], [3, 5, 1, 7], [2, 1, 0, 4]):
        b = prop.point(L, t_srce, 2, 1, device="cpu").numpy()
        lex = indep.fermion_to_lexico(b, L)
        x, y, z, t = t_srce
        assert lex[t, z, y, x, 2, 1] == 1 and np.count_nonzero(lex) == 1


def test_wall_and_momentum_sources():
    b = prop.wall(L, 5, 0, 2, device="cpu").numpy()
    lex = indep.fermion_to_lexico(b, L)
    assert np.count_nonzero(lex) == L[0] * L[1] * L[2]
    assert np.all(lex[5, :, :, :, 0, 2] == 1)
    phase = np.exp(1j * np.arange(np.prod(L)).reshape(2, L[3], L[2], L[1], L[0] // 2))
    m = prop.momentum(L, 3, phase, 1, 1, device="cpu").numpy()
    assert np.allclose(m[:, 3, :, :, :, 1, 1], phase[:, 3]) and np.count_nonzero(m) == L[0] * L[1] * L[2]


def test_propagator_shape_and_errors():
    assert prop.propagator_shape(L) == (2, 8, 2, 6, 2, 4, 4, 3, 3)
    try:
        prop.source(L, "gaussian", 0, 0, 0, device="cpu")
    except ValueError as e:
        assert "not supported" in str(e)
    else:
        raise AssertionError("unsupported source type accepted")


class _FakeLattice:
    """stands in for a Lattice: returns b as the solution with a chosen convergence flag"""
    kappa = 0.125

    def __init__(self, converged):
        self.converged = converged
        self.calls = []

    def solve(self, b, **kw):
        from qcu_amd.lattice import SolveResult
        self.calls.append(kw)
        return b.clone(), SolveResult(3, 0, self.converged, False, 1e-11, 0.0, 0)


def test_general_invert_normalisation_and_failure():
    from qcu_amd import general
    b = torch.ones((2, 2, 2, 2, 1, 4, 3), dtype=torch.complex128)
    lat = _FakeLattice(True)
    y, info = general.invert(b, lat, method="cgnr")
    assert torch.allclose(y, b * 0.25) and info.iters == 3 and lat.calls[0]["method"] == "cgnr"
    y_pc, _ = general.invertPC(b, lat)
    assert torch.equal(y_pc, y)
    try:
        general.invert(b, _FakeLattice(False))
    except RuntimeError as e:
        assert "did not reach" in str(e)
    else:
        raise AssertionError("a solve that did not converge was accepted")
