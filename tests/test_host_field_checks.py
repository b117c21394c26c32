"""The Python front end refuses fields the library would read past or
misinterpret: every method that hands a tensor to libqcu.so checks that it is
complex128 with exactly the element count of the field it stands for (a full
lattice, one parity, a gauge field), before any library call.  CPU only: the
handles are made without a device and never reach the library.
"""
import pytest
import torch

from qcu_amd import lattice as lt


def _lattice(L=(4, 4, 4, 8)):
    lat = object.__new__(lt.Lattice)  # no handle: the checks come first
    lat.L, lat.kappa, lat.csw, lat.t_boundary, lat.h = L, 0.12, 1.0, -1, None
    return lat


def _net(lat):
    net = object.__new__(lt.LearnedPreconditioner)
    net.lattice, net.channels, net.handle, net.nweights = lat, (1, 8, 1), None, 0
    return net


def test_wrong_dtype_and_size_are_refused():
    lat = _lattice()
    full, half = lt.fermion_shape(lat.L), lt.fermion_shape(lat.L, full=False)
    bad = [torch.zeros(full, dtype=torch.complex64),       # FP32 complex
           torch.zeros(half, dtype=torch.complex128),      # one parity where a full field is due
           torch.zeros(full, dtype=torch.float64)]
    for b in bad:
        with pytest.raises(ValueError):
            lat.solve(b)
        with pytest.raises(ValueError):
            lat.mat(b)
        with pytest.raises(ValueError):
            lat.true_residual(b, b)
    with pytest.raises(ValueError):
        lat.dslash(torch.zeros(full, dtype=torch.complex128), parity=0)  # full where one parity is due
    with pytest.raises(ValueError):
        lat.load_gauge(torch.zeros(lt.gauge_shape(lat.L), dtype=torch.complex64))
    with pytest.raises(ValueError):
        lat.solve(torch.zeros(full, dtype=torch.complex128), x=torch.zeros(half, dtype=torch.complex128))
    net = _net(lat)
    with pytest.raises(ValueError):
        net.apply_odd(torch.zeros(full, dtype=torch.complex128))
    with pytest.raises(ValueError):
        net.forward(torch.zeros(half, dtype=torch.complex128))
    with pytest.raises(ValueError):
        net.backward(torch.zeros((1,) + full, dtype=torch.complex128), torch.zeros(half, dtype=torch.complex128))
    with pytest.raises(ValueError):
        net.train([torch.zeros(half, dtype=torch.complex64)])
    with pytest.raises(ValueError):
        net.loss([torch.zeros(full, dtype=torch.complex128)])
    with pytest.raises(ValueError):
        net.loss_grad(torch.zeros(full, dtype=torch.complex128))
    with pytest.raises(ValueError):
        net.train_pairs([torch.zeros(half, dtype=torch.complex128)], [torch.zeros(full, dtype=torch.complex128)])


def test_batch_checks_each_system_against_its_lattice():
    a, b = _lattice((4, 4, 4, 8)), _lattice((4, 4, 4, 4))
    batch = object.__new__(lt.Batch)
    batch.lattices, batch.nrhs, batch.nsys, batch.h = [a, b], [2, 1], 3, None
    fa = torch.zeros(lt.fermion_shape(a.L), dtype=torch.complex128)
    fb = torch.zeros(lt.fermion_shape(b.L), dtype=torch.complex128)
    with pytest.raises(ValueError):
        batch.solve([fa, fa])             # one right-hand side short
    with pytest.raises(ValueError):
        batch.solve([fa, fb, fb])         # system 1 belongs to lattice a
    with pytest.raises(ValueError):
        batch.solve([fa, fa, fb.to(torch.complex64)])
