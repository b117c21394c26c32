"""Pins the oracle's learned-preconditioner network (forward and backward).

The network is new (no counterpart in the reference), so its oracle is pinned
by an independent numpy formulation (np.roll shifts, einsum spin and channel
mixing), by gauge covariance, by linearity, by the adjoint identity
<g, K x> = <K^+ g, x> for the input gradient, and by central finite
differences for the weight gradient.
"""
import numpy as np
import pytest

import indep

CASES = [((4, 4, 2, 4), (1, 3, 2, 1), 1), ((2, 2, 4, 6), (1, 2, 1), -1), ((4, 2, 2, 2), (1, 1), -1)]


def _weights(orc, chan, seed):
    n = orc.net_weight_count(chan)
    return orc.gauss((n,), seed) * 0.4


def _split(W, chan):
    out, off = [], 0
    for ci, co in zip(chan[:-1], chan[1:]):
        k = 9 * co * ci * 16
        out.append(W[off:off + k].reshape(9, co, ci, 4, 4))
        off += k
    return out


def _fields_lex(v, c, L):
    return np.stack([indep.fermion_to_lexico(v.reshape(c, -1)[k], L) for k in range(c)])


@pytest.mark.parametrize("L,chan,tb", CASES)
def test_net_forward_matches_independent(orc, L, chan, tb):
    U = orc.gauge_random(L, 0.5, 11)
    W = _weights(orc, chan, 12)
    x = orc.gauss((chan[0], 2, orc.vol(L) // 2, 4, 3), 13)
    out, _ = orc.net_forward(x, W, chan, U, L, tb)
    ref = indep.net_lexico(_fields_lex(x, chan[0], L), _split(W, chan), indep.gauge_to_lexico(U, L), L, tb)
    got = _fields_lex(out, chan[-1], L)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-13


def test_net_identity_weights(orc):
    """C = 1 on channel 0 and everything else zero gives the identity"""
    L, chan = (4, 4, 4, 4), (1, 2, 1)
    U = orc.gauge_random(L, 0.5, 21)
    Ws = [np.zeros((9, co, ci, 4, 4), complex) for ci, co in zip(chan[:-1], chan[1:])]
    for W in Ws:
        W[8, 0, 0] = np.eye(4)
    x = orc.gauss((1, 2, orc.vol(L) // 2, 4, 3), 22)
    out, _ = orc.net_forward(x, np.concatenate([w.ravel() for w in Ws]), chan, U, L, -1)
    assert np.array_equal(out.reshape(x.shape), x)


def test_net_gauge_covariance(orc):
    L, chan, tb = (4, 4, 4, 4), (1, 2, 2, 1), -1
    rng = np.random.default_rng(5)
    U = orc.gauge_random(L, 0.5, 31)
    W = _weights(orc, chan, 32)
    x = orc.gauss((1, 2, orc.vol(L) // 2, 4, 3), 33)
    Ul = indep.gauge_to_lexico(U, L)
    g = indep.random_su3_gauge_transform(L, rng)
    Ug = indep.gauge_from_lexico(indep.gauge_transform(Ul, g), L)
    gx = indep.fermion_to_cb(np.einsum("tzyxab,tzyxsb->tzyxsa", g, indep.fermion_to_lexico(x[0], L)), L)
    out, _ = orc.net_forward(x, W, chan, U, L, tb)
    outg, _ = orc.net_forward(gx[None], W, chan, Ug, L, tb)
    want = np.einsum("tzyxab,tzyxsb->tzyxsa", g, indep.fermion_to_lexico(out, L))
    got = indep.fermion_to_lexico(outg, L)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-13


def test_net_linear_in_input(orc):
    L, chan = (4, 4, 2, 4), (1, 3, 1)
    U = orc.gauge_random(L, 0.5, 41)
    W = _weights(orc, chan, 42)
    x = orc.gauss((1, 2, orc.vol(L) // 2, 4, 3), 43)
    y = orc.gauss((1, 2, orc.vol(L) // 2, 4, 3), 44)
    a, b = 0.3 - 1.1j, -2.0 + 0.5j
    kx, _ = orc.net_forward(x, W, chan, U, L, 1)
    ky, _ = orc.net_forward(y, W, chan, U, L, 1)
    kxy, _ = orc.net_forward(a * x + b * y, W, chan, U, L, 1)
    assert np.linalg.norm(kxy - a * kx - b * ky) / np.linalg.norm(kxy) < 1e-14


@pytest.mark.parametrize("L,chan,tb", CASES)
def test_net_input_gradient_is_adjoint(orc, L, chan, tb):
    """grad_in for grad_out = g is K^+ g: <g, K x> = <K^+ g, x>"""
    U = orc.gauge_random(L, 0.5, 51)
    W = _weights(orc, chan, 52)
    x = orc.gauss((chan[0], 2, orc.vol(L) // 2, 4, 3), 53)
    g = orc.gauss((chan[-1], 2, orc.vol(L) // 2, 4, 3), 54)
    kx, acts = orc.net_forward(x, W, chan, U, L, tb)
    kg, _ = orc.net_backward(g, W, chan, U, L, acts, tb)
    lhs = np.vdot(g.ravel(), kx)
    rhs = np.vdot(kg, x.ravel())
    assert abs(lhs - rhs) / abs(lhs) < 1e-13


@pytest.mark.parametrize("L,chan,tb", CASES[:2])
def test_net_weight_gradient_finite_differences(orc, L, chan, tb):
    """L(W) = |K_W x - b|^2: directional derivatives from central differences"""
    U = orc.gauge_random(L, 0.5, 61)
    W = _weights(orc, chan, 62)
    x = orc.gauss((chan[0], 2, orc.vol(L) // 2, 4, 3), 63)
    b = orc.gauss((chan[-1], 2, orc.vol(L) // 2, 4, 3), 64).ravel()

    def loss(Wv):
        out, acts = orc.net_forward(x, Wv, chan, U, L, tb)
        return float(np.vdot(out - b, out - b).real), out, acts

    l0, out, acts = loss(W)
    _, gW = orc.net_backward(2 * (out - b), W, chan, U, L, acts, tb)
    rng = np.random.default_rng(7)
    eps = 1e-5
    for _ in range(3):
        d = rng.normal(size=W.shape) + 1j * rng.normal(size=W.shape)
        fd = (loss(W + eps * d)[0] - loss(W - eps * d)[0]) / (2 * eps)
        an = float(np.vdot(gW, d).real)
        assert abs(fd - an) <= 1e-6 * abs(an) + 1e-9 * l0, (fd, an)
    # a single weight: the C term of the last layer, one spin entry
    e = np.zeros_like(W)
    e[-1] = 1.0
    fd = (loss(W + eps * e)[0] - loss(W - eps * e)[0]) / (2 * eps)
    assert abs(fd - gW[-1].real) <= 1e-6 * abs(gW[-1].real) + 1e-9 * l0
