"""GPU parity of the learned preconditioner's HIP kernels against the oracle.

Forward, input gradient and weight gradient are compared with the oracle's
network (itself pinned in test_oracle_net.py) on the same FP32-rounded
weights; the training loss gradient is compared with the oracle chain
K -> Mhat -> |.|^2 whose Mhat^+ is pinned here by the adjoint identity.
The last test trains on the GPU and shows the iteration gain of the
preconditioned solve, every solve checked by the oracle's FP64 true residual.
FP32 tolerances are stated in each test.
"""
import numpy as np
import pytest

from gpu_util import dev, host, log, rel

pytestmark = pytest.mark.gpu

G5 = np.array([1, 1, -1, -1], float)


def _w32(w):
    return w.astype(np.complex64).astype(np.complex128)


def _setup(orc, L, chan, tb, seed, csw=0.0, kappa=0.124, eps=0.4):
    from qcu_amd.lattice import Lattice, LearnedPreconditioner
    U = orc.gauge_random(L, eps, seed)
    lat = Lattice(L, kappa, csw, tb)
    lat.load_gauge(dev(U))
    net = LearnedPreconditioner(lat, chan, seed=seed)
    W = _w32(orc.gauss((net.nweights,), seed + 1) * 0.3)
    net.set_weights(W)
    return U, lat, net, W


# Lx / 2 a multiple of 16 selects the tiled wave order of the matrix-core layer
# kernel (tiles of 8 x 6 and 2 x 4 rows here); the smaller lattices run it in
# site order.  (6, 2, 2, 2) has 24 sites per parity: the matrix-core kernels
# (16 sites per wave) end on a partial wave there.
@pytest.mark.parametrize("L,chan,tb", [((8, 4, 4, 8), (1, 8, 8, 8, 1), -1), ((4, 6, 4, 4), (1, 2, 3, 1), 1),
                                       ((6, 4, 6, 2), (1, 1), -1), ((32, 8, 6, 4), (1, 8, 5, 8, 1), -1),
                                       ((32, 2, 4, 2), (1, 3, 1), 1), ((6, 2, 2, 2), (1, 8, 3, 1), 1)])
def test_net_forward_and_backward_vs_oracle(orc, L, chan, tb):
    U, lat, net, W = _setup(orc, L, chan, tb, 100)
    assert net.nweights == orc.net_weight_count(chan)
    x = orc.gauss((1, 2, orc.vol(L) // 2, 4, 3), 101)
    g = orc.gauss((1, 2, orc.vol(L) // 2, 4, 3), 102)
    ref, acts = orc.net_forward(x, W, chan, U, L, tb)
    out = host(net.forward(dev(x)))
    e_fwd = rel(out.ravel(), ref)
    gin_ref, gW_ref = orc.net_backward(g, W, chan, U, L, acts, tb)
    gW, gin = net.backward(dev(x), dev(g))
    e_gin = rel(host(gin).ravel(), gin_ref)
    e_gw = rel(gW, gW_ref)
    log(f"L={L} chan={chan} tb={tb}: forward {e_fwd:.1e}, input grad {e_gin:.1e}, weight grad {e_gw:.1e}")
    # FP32 activations and weights, sums of at most 9*8 terms per layer
    assert e_fwd < 2e-5 and e_gin < 2e-5
    # weight gradient: FP32 site products, blockwise FP32 sums, FP64 across blocks
    assert e_gw < 1e-4


# apply_odd takes the split first layer when the first layer is 1 -> 8
# (followed here by no split layer, by an 8 -> 4 split layer, or by nothing
# but the last layer), the FP32 kernels for 1 -> 4.
@pytest.mark.parametrize("chan", [(1, 4, 4, 1), (1, 8, 1), (1, 8, 4, 1)])
@pytest.mark.parametrize("L", [(8, 8, 4, 4), (32, 4, 4, 4), (6, 2, 2, 2)])
def test_net_apply_odd_is_odd_part_of_forward(orc, L, chan):
    U, lat, net, W = _setup(orc, L, chan, -1, 110)
    xo = orc.gauss((orc.vol(L) // 2, 4, 3), 111)
    full = np.zeros((1, 2) + xo.shape, complex)
    full[0, 1] = xo
    ref, _ = orc.net_forward(full, W, chan, U, L, -1)
    got = host(net.apply_odd(dev(xo)))
    e = rel(got.ravel(), ref.reshape(2, -1)[1])
    log(f"apply_odd chan={chan} L={L} vs oracle {e:.1e}")
    assert e < 2e-5


def _schur_ops(orc, U, L, kappa, csw, tb):
    """Mhat and Mhat^+ on odd vectors from oracle pieces"""
    coeff = kappa * csw
    if csw:
        Ao_inv = orc.clover_invert(orc.clover_build(U, L, 1, coeff))
    else:
        Ao_inv = None

    def mhat(y):
        t = orc.clover_dslash(y, U, L, 0, coeff, tb)
        return y - kappa ** 2 * orc.clover_dslash(t, U, L, 1, coeff, tb)

    def mhat_dag(r):
        z = r if Ao_inv is None else orc.clover_apply(Ao_inv, r)
        z = z * G5[None, :, None]
        t = orc.clover_dslash(z, U, L, 0, coeff, tb)
        t = orc.hop(t, U, L, 1, tb)
        return r - kappa ** 2 * t * G5[None, :, None]

    return mhat, mhat_dag


# In training the network input and the output gradient are zero on the even
# sites, and the first and last layers' weight gradients sweep only the
# parities that can contribute; (6, 2, 2, 2) (24 sites per parity) ends those
# sweeps on a partial 4-site group and a partial 16-site wave; (16, 8, 16, 16)
# runs the first layer's kernel in its z-slab chunk order.
@pytest.mark.parametrize("L,chan,csw", [((8, 4, 4, 8), (1, 4, 4, 1), 0.0), ((8, 4, 4, 8), (1, 4, 4, 1), 1.0),
                                        ((6, 2, 2, 2), (1, 8, 3, 1), 1.0), ((16, 8, 16, 16), (1, 8, 8, 8, 1), 1.0)])
def test_training_gradient_vs_oracle(orc, L, chan, csw):
    tb, kappa = -1, 0.124
    U, lat, net, W = _setup(orc, L, chan, tb, 120, csw=csw, kappa=kappa)
    net.set_weights(W * 0.2)  # a smaller net keeps Mhat K b near b
    W = net.get_weights()
    mhat, mhat_dag = _schur_ops(orc, U, L, kappa, csw, tb)
    vh = orc.vol(L) // 2
    u = orc.gauss((vh, 4, 3), 121)
    v = orc.gauss((vh, 4, 3), 122)
    adj = abs(np.vdot(u, mhat(v)) - np.vdot(mhat_dag(u), v)) / abs(np.vdot(u, mhat(v)))
    assert adj < 1e-13  # pins the Mhat^+ used below
    b = _w32(orc.gauss((vh, 4, 3), 123))
    full = np.zeros((1, 2, vh, 4, 3), complex)
    full[0, 1] = b
    y, acts = orc.net_forward(full, W, chan, U, L, tb)
    y_odd = y.reshape(2, vh, 4, 3)[1]
    res = mhat(y_odd) - b
    b2 = np.vdot(b, b).real
    loss_ref = np.vdot(res, res).real / b2
    gy = np.zeros((1, 2, vh, 4, 3), complex)
    gy[0, 1] = 2 * mhat_dag(res) / b2
    _, gW_ref = orc.net_backward(gy, W, chan, U, L, acts, tb)
    loss, gW = net.loss_grad(dev(b))
    e_l, e_g = abs(loss - loss_ref) / loss_ref, rel(gW, gW_ref)
    log(f"L={L} chan={chan} csw={csw}: loss {loss:.6f} (oracle {loss_ref:.6f}, rel {e_l:.1e}), "
        f"gradient rel {e_g:.1e}")
    # FP32 operator and network; the loss is a difference of O(1) terms
    assert e_l < 1e-4 and e_g < 1e-3


# The preconditioned inner solve in each precision: the last layer's gather
# writes the network output straight into the solver's vector of that precision.
@pytest.mark.parametrize("prec_name", ["fp32", "fp64", "fp16"])
def test_trained_preconditioner_reduces_iterations(orc, prec_name):
    """Adam on the GPU lowers the loss; the preconditioned solve needs fewer
    iterations than the plain one and still meets |b - M x|/|b| <= 1e-10"""
    from qcu_amd.lattice import PREC, Lattice, LearnedPreconditioner
    prec = PREC[prec_name]
    L, kappa, csw, tb = (8, 8, 8, 16), 0.1245, 1.0, -1
    U = orc.gauge_random(L, 0.25, 130)
    lat = Lattice(L, kappa, csw, tb)
    lat.load_gauge(dev(U))
    vh = orc.vol(L) // 2
    srcs = [dev(orc.gauss((vh, 4, 3), 1000 + k)) for k in range(32)]
    net = LearnedPreconditioner(lat, (1, 8, 8, 8, 1), seed=3)
    l0 = net.loss(srcs)
    losses, secs = net.train(srcs, epochs=25, batch=4, lr=2e-3)
    l1 = net.loss(srcs)
    log(f"loss {l0:.4f} -> {l1:.4f} in {secs:.1f} s; per epoch {np.round(losses, 4).tolist()}")
    assert l1 < 0.8 * l0
    b = dev(orc.gauss((2, vh, 4, 3), 140))
    x0, r0 = lat.solve(b, tol=1e-10, prec=prec)
    x1, r1 = lat.solve(b, tol=1e-10, prec=prec, precond=net)
    t0 = orc.true_residual(host(x0), host(b), U, L, kappa, csw, tb)
    t1 = orc.true_residual(host(x1), host(b), U, L, kappa, csw, tb)
    log(f"{prec_name}: plain {r0.iters} iters ({r0.seconds * 1e3:.1f} ms, true res {t0:.1e}); "
        f"learned {r1.iters} iters ({r1.seconds * 1e3:.1f} ms, true res {t1:.1e}, fallback {r1.fallback})")
    assert r0.converged and r1.converged and t0 <= 1e-10 and t1 <= 1e-10
    assert not r1.fallback
    assert r1.iters < r0.iters


@pytest.mark.parametrize("fill", [0.0, np.nan])
def test_broken_preconditioner_falls_back_to_plain(orc, fill):
    """a network whose output is zero (or NaN) breaks the preconditioned
    BiCGSTAB down at its first step (<r~, Mhat K p> is 0, or not finite); after
    the bounded restarts the solve continues without the network from the
    current solution and still meets |b - M x|/|b| <= 1e-10"""
    from qcu_amd.lattice import FP16, Lattice, LearnedPreconditioner
    L, kappa, csw, tb = (8, 4, 4, 8), 0.124, 1.0, -1
    U = orc.gauge_random(L, 0.3, 145)
    lat = Lattice(L, kappa, csw, tb)
    lat.load_gauge(dev(U))
    net = LearnedPreconditioner(lat, (1, 4, 1), seed=146)
    net.set_weights(np.full(net.nweights, complex(fill, fill)))
    b = orc.gauss((2, orc.vol(L) // 2, 4, 3), 147)
    x, r = lat.solve(dev(b), tol=1e-10, prec=FP16, reliable_delta=0.01, precond=net)
    tr = orc.true_residual(host(x), b, U, L, kappa, csw, tb)
    log(f"weights {fill}: {r.iters} iters, {r.restarts} restarts, fallback {r.fallback}, oracle true residual {tr:.1e}")
    assert r.fallback and r.converged and tr <= 1e-10
    net.close()
    lat.close()


def test_vector_layer_kernel_matches_oracle(orc, monkeypatch):
    """the vector-FMA layer kernel (kept as a cross-check of the matrix-core one)"""
    monkeypatch.setenv("QCU_NET_KERNEL", "valu")
    L, chan, tb = (8, 4, 4, 8), (1, 8, 8, 8, 1), -1
    U, lat, net, W = _setup(orc, L, chan, tb, 150)
    x = orc.gauss((1, 2, orc.vol(L) // 2, 4, 3), 151)
    ref, _ = orc.net_forward(x, W, chan, U, L, tb)
    e = rel(host(net.forward(dev(x))).ravel(), ref)
    log(f"vector layer kernel: forward {e:.1e}")
    assert e < 2e-5


# Hidden 8 -> 8 and 8 -> 4 layers run on the FP16 matrix cores with split
# operands (x = hi + lo, three FP16 products accumulated in FP32).  (6, 2, 2, 2)
# has 24 sites per parity and (10, 6, 4, 4) 240: both end on a partial 32-site
# chunk; order 1 deals one eighth of the chunks to each XCD.
@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("L,chan", [((8, 4, 4, 8), (1, 8, 8, 8, 1)), ((6, 2, 2, 2), (1, 8, 8, 4, 1)),
                                    ((10, 6, 4, 4), (1, 8, 8, 1))])
def test_split_fp16_layers_match_oracle_and_fp32_kernel(orc, monkeypatch, L, chan, order):
    """forward and input gradient with the split-FP16 hidden layers against the
    oracle (FP32 bound 2e-5) and against the FP32 matrix-core kernel (2e-6)"""
    tb = -1
    monkeypatch.setenv("QCU_NET_SPLIT_ORDER", str(order))
    U, lat, net, W = _setup(orc, L, chan, tb, 170)
    monkeypatch.setenv("QCU_NET_SPLIT", "0")
    from qcu_amd.lattice import LearnedPreconditioner
    ref32 = LearnedPreconditioner(lat, chan, seed=170)
    ref32.set_weights(W)
    x = orc.gauss((1, 2, orc.vol(L) // 2, 4, 3), 171)
    g = orc.gauss((1, 2, orc.vol(L) // 2, 4, 3), 172)
    ref, acts = orc.net_forward(x, W, chan, U, L, tb)
    gin_ref, gW_ref = orc.net_backward(g, W, chan, U, L, acts, tb)
    out = host(net.forward(dev(x))).ravel()
    out32 = host(ref32.forward(dev(x))).ravel()
    gW, gin = net.backward(dev(x), dev(g))
    gW32, gin32 = ref32.backward(dev(x), dev(g))
    gin, gin32 = host(gin).ravel(), host(gin32).ravel()
    e = (rel(out, ref), rel(out, out32), rel(gin, gin_ref), rel(gin, gin32), rel(gW, gW_ref))
    log(f"split L={L} chan={chan} order {order}: forward {e[0]:.1e} (FP32 kernel {e[1]:.1e}), "
        f"input grad {e[2]:.1e} ({e[3]:.1e}), weight grad {e[4]:.1e}")
    assert e[0] < 2e-5 and e[2] < 2e-5 and e[4] < 1e-4
    assert e[1] < 2e-6 and e[3] < 2e-6
    ref32.close()
    net.close()


# A wave walks several chunks when the lattice has more 32-site chunks than
# the grid has waves (256 CUs x 8): 16x16x18x32 has 2 x 36864 / 32 = 4608, so
# waves take 2 or 3 chunks, and each chunk's first loads are issued while the
# previous chunk's output is stored.  Forward (plain epilogue) and apply_odd
# (the fused epilogue) against the oracle and the FP32 matrix-core kernel.
@pytest.mark.parametrize("order", [0, 1])
def test_split_layers_several_chunks_per_wave(orc, monkeypatch, order):
    """split-FP16 layers on a lattice with 2-3 chunks per wave"""
    L, chan, tb = (16, 16, 18, 32), (1, 8, 8, 8, 1), -1
    monkeypatch.setenv("QCU_NET_SPLIT_ORDER", str(order))
    U, lat, net, W = _setup(orc, L, chan, tb, 210)
    monkeypatch.setenv("QCU_NET_SPLIT", "0")
    from qcu_amd.lattice import LearnedPreconditioner
    ref32 = LearnedPreconditioner(lat, chan, seed=210)
    ref32.set_weights(W)
    x = orc.gauss((1, 2, orc.vol(L) // 2, 4, 3), 211)
    ref, _ = orc.net_forward(x, W, chan, U, L, tb)
    out = host(net.forward(dev(x))).ravel()
    out32 = host(ref32.forward(dev(x))).ravel()
    xo = np.ascontiguousarray(x[0, 1])
    full = np.zeros_like(x)
    full[0, 1] = xo
    ref_odd, _ = orc.net_forward(full, W, chan, U, L, tb)
    got_odd = host(net.apply_odd(dev(xo))).ravel()
    got_odd32 = host(ref32.apply_odd(dev(xo))).ravel()
    e = (rel(out, ref), rel(out, out32), rel(got_odd, ref_odd.reshape(2, -1)[1]), rel(got_odd, got_odd32))
    log(f"split L={L} order {order}: forward {e[0]:.1e} (FP32 kernel {e[1]:.1e}), "
        f"apply_odd {e[2]:.1e} ({e[3]:.1e})")
    assert e[0] < 2e-5 and e[2] < 2e-5
    assert e[1] < 2e-6 and e[3] < 2e-6
    ref32.close()
    net.close()


# Both MFMA shapes of the split kernel (32 sites per wave on 32x32x16, 16 on
# 16x16x32) and its chunk orders: 2 walks a z slab per XCD; 3 (16 sites, Lx
# 32, Lz a multiple of 16) gives each wave one x row and rotates its term
# order by (y + 2 z + 3 t) mod 7.
@pytest.mark.parametrize("shape,order", [(32, 2), (16, 0), (16, 2), (16, 3)])
def test_split_shapes_and_orders(orc, monkeypatch, shape, order):
    """forward and input gradient of a 1-8-8-8-1 network on 32x2x16x2 against
    the oracle (2e-5) for each split shape and chunk order"""
    monkeypatch.setenv("QCU_NET_SPLIT_SHAPE", str(shape))
    monkeypatch.setenv("QCU_NET_SPLIT_ORDER", str(order))
    L, chan, tb = (32, 2, 16, 2), (1, 8, 8, 8, 1), -1
    U, lat, net, W = _setup(orc, L, chan, tb, 200)
    x = orc.gauss((1, 2, orc.vol(L) // 2, 4, 3), 201)
    g = orc.gauss((1, 2, orc.vol(L) // 2, 4, 3), 202)
    ref, acts = orc.net_forward(x, W, chan, U, L, tb)
    gin_ref, _ = orc.net_backward(g, W, chan, U, L, acts, tb)
    out = host(net.forward(dev(x))).ravel()
    _, gin = net.backward(dev(x), dev(g))
    xo = x[0, 1]
    full = np.zeros_like(x)
    full[0, 1] = xo
    ref_odd, _ = orc.net_forward(full, W, chan, U, L, tb)
    got_odd = host(net.apply_odd(dev(np.ascontiguousarray(xo)))).ravel()
    e = (rel(out, ref), rel(host(gin).ravel(), gin_ref), rel(got_odd, ref_odd.reshape(2, -1)[1]))
    log(f"split shape {shape} order {order}: forward {e[0]:.1e}, input grad {e[1]:.1e}, apply_odd {e[2]:.1e}")
    assert max(e) < 2e-5
    net.close()


# The preconditioner application folds the last layer (8 -> 1) into the
# epilogue of the 8 -> 8 layer before it: that kernel writes the last layer's
# mixed term fields instead of its own output channels.
# (16, 8, 16, 16) is large enough for the z-slab chunk order (a grid of a
# multiple of 8 blocks, Lz a multiple of 8) of the second and third layers.
@pytest.mark.parametrize("L,chan", [((8, 4, 4, 8), (1, 8, 8, 8, 1)), ((6, 2, 2, 2), (1, 8, 8, 1)),
                                    ((10, 6, 4, 4), (1, 8, 8, 8, 1)), ((16, 8, 16, 16), (1, 8, 8, 8, 1))])
def test_fused_last_layer_apply_odd(orc, monkeypatch, L, chan):
    """apply_odd with the fused last layer against the oracle (FP32 bound 2e-5)
    and against the unfused FP32 matrix-core path (2e-6)"""
    U, lat, net, W = _setup(orc, L, chan, -1, 190)
    monkeypatch.setenv("QCU_NET_SPLIT", "0")
    from qcu_amd.lattice import LearnedPreconditioner
    ref32 = LearnedPreconditioner(lat, chan, seed=190)
    ref32.set_weights(W)
    xo = orc.gauss((orc.vol(L) // 2, 4, 3), 191)
    full = np.zeros((1, 2) + xo.shape, complex)
    full[0, 1] = xo
    ref, _ = orc.net_forward(full, W, chan, U, L, -1)
    ref = ref.reshape(2, -1)[1]
    got = host(net.apply_odd(dev(xo))).ravel()
    got32 = host(ref32.apply_odd(dev(xo))).ravel()
    # the full forward (all activations kept, no fusion) agrees too
    fwd = host(net.forward(dev(full))).reshape(2, -1)[1]
    e = (rel(got, ref), rel(got, got32), rel(fwd, ref))
    log(f"fused apply_odd L={L} chan={chan}: vs oracle {e[0]:.1e}, vs FP32 kernels {e[1]:.1e}, "
        f"unfused forward vs oracle {e[2]:.1e}")
    assert e[0] < 2e-5 and e[1] < 2e-6 and e[2] < 2e-5
    ref32.close()
    net.close()


@pytest.mark.parametrize("scale", [1e-25, 1e-8, 1e10, 1e25])
def test_split_layers_follow_the_input_scale(orc, scale):
    """K is linear: K(s x) / s = K(x) to FP32 accuracy for inputs far from unit
    size (FP16 alone spans 6e-8 .. 6.5e4; the split kernel scales its operands
    by powers of two from the largest component the previous layer wrote)"""
    L, chan, tb = (8, 4, 4, 8), (1, 8, 8, 8, 1), -1
    U, lat, net, W = _setup(orc, L, chan, tb, 180)
    x = orc.gauss((1, 2, orc.vol(L) // 2, 4, 3), 181)
    ref, _ = orc.net_forward(x, W, chan, U, L, tb)
    out1 = host(net.forward(dev(x))).ravel()
    outs = host(net.forward(dev(x * scale))).ravel() / scale
    e_ref, e_one = rel(outs, ref), rel(outs, out1)
    log(f"input scale {scale:g}: K(s x)/s vs oracle K(x) {e_ref:.1e}, vs K(x) {e_one:.1e}")
    assert e_ref < 2e-5 and e_one < 2e-6
    net.close()


# The preconditioner's first two layers run as net_layer_first on the even
# sites and net_layer_second, which forms the second layer's even outputs from
# the input with combined matrices W2_t W1_8 + W2_8 W1_t and takes its odd
# outputs' site term from the input too, rescaling its accumulators between
# the input's scale and the first layer's.  The third layer takes its site
# term at its even output sites from the input as well (W3_8 times those
# combined matrices), rescaling between the input's scale and the second
# layer's.  Scaling W1 by a and W2 by 1/a leaves the network unchanged while
# moving the first layer's output far from the input's size; W2 by b and W3
# by 1/b does the same for the second layer's output; the input itself is
# also scaled.
@pytest.mark.parametrize("a,b,scale", [(1.0, 1.0, 1.0), (1e-12, 1.0, 1.0), (1e12, 1.0, 1.0), (1.0, 1.0, 1e-20),
                                       (1e-6, 1.0, 1e15), (1.0, 1e-12, 1.0), (1.0, 1e12, 1e-10),
                                       (1e6, 1e-9, 1e12)])
def test_second_layer_scales(orc, a, b, scale):
    """apply_odd of a 1-8-8-8-1 network with W1 * a, W2 * b / a, W3 / b and
    input * scale against the oracle's K(x) (FP32 bound 2e-5)"""
    L, chan, tb = (8, 4, 4, 8), (1, 8, 8, 8, 1), -1
    U, lat, net, W = _setup(orc, L, chan, tb, 220)
    n0, n1 = 9 * 1 * 8 * 16, 9 * 8 * 8 * 16
    Ws = W.copy()
    Ws[:n0] *= a
    Ws[n0:n0 + n1] *= b / a
    Ws[n0 + n1:n0 + 2 * n1] /= b
    net.set_weights(Ws)
    xo = orc.gauss((orc.vol(L) // 2, 4, 3), 221)
    full = np.zeros((1, 2) + xo.shape, complex)
    full[0, 1] = xo
    ref, _ = orc.net_forward(full, W, chan, U, L, tb)
    got = host(net.apply_odd(dev(xo * scale))).ravel() / scale
    e = rel(got, ref.reshape(2, -1)[1])
    log(f"W1 x {a:g}, W2 x {b:g} / {a:g}, W3 / {b:g}, input x {scale:g}: apply_odd vs oracle {e:.1e}")
    assert e < 2e-5
    net.close()


# The 8 -> 8 weight gradients run on the FP16 matrix cores with split
# operands, scaled by powers of two from the recorded field maxima; the
# gradient is bilinear in (input, output gradient).
# Its operands reach the matrix cores through LDS (default) or by direct
# loads (QCU_NET_WGRAD_LDS=0); (6, 2, 2, 2) has 48 sites (both parities), so
# the second 32-site step of the LDS path is partial.
@pytest.mark.parametrize("lds", ["1", "0"])
@pytest.mark.parametrize("sx,sg,L", [(1.0, 1.0, (8, 4, 4, 8)), (1e-15, 1e15, (8, 4, 4, 8)),
                                     (1e8, 1e-12, (8, 4, 4, 8)), (1.0, 1.0, (6, 2, 2, 2))])
def test_split_weight_gradient(orc, monkeypatch, sx, sg, L, lds):
    """weight gradient with split-FP16 8 -> 8 products: gW(sx x, sg g) / (sx sg)
    against the oracle's gW(x, g) (1e-4, as the FP32 kernels) and against the
    FP32 matrix-core weight gradient (1e-5)"""
    monkeypatch.setenv("QCU_NET_WGRAD_LDS", lds)
    chan, tb = (1, 8, 8, 8, 1), -1
    U, lat, net, W = _setup(orc, L, chan, tb, 210)
    monkeypatch.setenv("QCU_NET_SPLIT_WGRAD", "0")
    from qcu_amd.lattice import LearnedPreconditioner
    ref32 = LearnedPreconditioner(lat, chan, seed=210)
    ref32.set_weights(W)
    x = orc.gauss((1, 2, orc.vol(L) // 2, 4, 3), 211)
    g = orc.gauss((1, 2, orc.vol(L) // 2, 4, 3), 212)
    _, acts = orc.net_forward(x, W, chan, U, L, tb)
    _, gW_ref = orc.net_backward(g, W, chan, U, L, acts, tb)
    gW, _ = net.backward(dev(x * sx), dev(g * sg), want_input_grad=False)
    gW = gW / (sx * sg)
    gW32, _ = ref32.backward(dev(x), dev(g), want_input_grad=False)
    e = (rel(gW, gW_ref), rel(gW, gW32))
    log(f"split weight gradient L={L} lds={lds}, scales {sx:g} {sg:g}: vs oracle {e[0]:.1e}, "
        f"vs FP32 kernel {e[1]:.1e}")
    assert e[0] < 1e-4 and e[1] < 1e-5
    ref32.close()
    net.close()


def test_replacing_the_network_recaptures_graphs(orc):
    """a solve with a new network (after the previous one was destroyed) must not
    replay graphs captured with the old one: graph and eager runs agree bit for bit"""
    import torch
    from qcu_amd.lattice import FP32, Lattice, LearnedPreconditioner
    L = (4, 4, 4, 8)
    U = orc.gauge_random(L, 0.3, 160)
    lat = Lattice(L, 0.124, 1.0, -1)
    lat.load_gauge(dev(U))
    b = dev(orc.gauss((2, orc.vol(L) // 2, 4, 3), 161))
    net1 = LearnedPreconditioner(lat, (1, 2, 1), seed=1, init_scale=0.3)
    x1, r1 = lat.solve(b, tol=1e-10, prec=FP32, precond=net1, use_graph=True)
    net1.close()
    net2 = LearnedPreconditioner(lat, (1, 2, 1), seed=2, init_scale=0.3)
    xg, rg = lat.solve(b, tol=1e-10, prec=FP32, precond=net2, use_graph=True)
    xe, re_ = lat.solve(b, tol=1e-10, prec=FP32, precond=net2, use_graph=False)
    log(f"net1 {r1.iters} iters; net2 graph {rg.iters} / eager {re_.iters} iters")
    assert rg.converged and torch.equal(xg, xe) and rg.iters == re_.iters


def test_schur_solutions_and_pair_training(orc):
    """(source, solution) pairs: schur_solutions gives Mhat y = c under the oracle's
    Mhat, and train_pairs lowers the pair loss |K c - y|^2 / |y|^2"""
    from qcu_amd.lattice import FP64, Lattice, LearnedPreconditioner, schur_solutions
    L, kappa, csw, tb = (8, 4, 4, 8), 0.124, 1.0, -1
    U = orc.gauge_random(L, 0.3, 130)
    lat = Lattice(L, kappa, csw, tb)
    lat.load_gauge(dev(U))
    cs = [dev(orc.gauss((orc.vol(L) // 2, 4, 3), 131 + k)) for k in range(8)]
    ys = schur_solutions(lat, cs, tol=1e-10, prec=FP64)
    mhat, _ = _schur_ops(orc, U, L, kappa, csw, tb)
    worst = max(rel(mhat(host(y)), host(c)) for c, y in zip(cs, ys))
    net = LearnedPreconditioner(lat, (1, 4, 4, 1), seed=132)
    # an epoch at learning rate 0 measures the pair loss at fixed weights
    before, _ = net.train_pairs(cs, ys, epochs=1, batch=4, lr=0.0)
    losses, secs = net.train_pairs(cs, ys, epochs=3, batch=4, lr=1e-4)
    after, _ = net.train_pairs(cs, ys, epochs=1, batch=4, lr=0.0)
    log(f"schur_solutions: max |Mhat y - c|/|c| {worst:.1e}; pair loss {before[0]:.4f} -> {after[0]:.4f} "
        f"(epoch means {[round(float(x), 4) for x in losses]})")
    assert worst < 1e-8
    assert np.isfinite(losses).all() and after[0] < before[0]
    net.close()
    lat.close()


def test_weights_file_round_trip(orc, tmp_path):
    """qcuNetSave / qcuNetLoad: the weights come back bit for bit (FP32), the
    loaded network applies the same K, the metadata line records the channel
    list and lattice, and a file for other channels or a truncated file is
    refused with ERR_ARG"""
    import json
    from qcu_amd._lib import QcuError
    from qcu_amd.lattice import Lattice, LearnedPreconditioner
    L, chan = (8, 4, 4, 8), (1, 8, 8, 8, 1)
    U, lat, net, W = _setup(orc, L, chan, -1, 140)
    path = tmp_path / "k.qcunet"
    net.save(path)
    raw = path.read_bytes()
    head, _, body = raw.partition(b"\n")
    assert head.startswith(b"QCUNET1 ")
    meta = json.loads(head[8:])
    assert meta["channels"] == list(chan) and meta["weights"] == net.nweights
    assert meta["lattice"] == list(L) and meta["dtype"] == "complex64-le"
    assert len(body) == 8 * net.nweights
    assert np.array_equal(np.frombuffer(body, "<c8").astype(np.complex128), W)
    # a second lattice of another shape loads the same weights
    L2 = (4, 4, 4, 8)
    U2 = orc.gauge_random(L2, 0.3, 141)
    lat2 = Lattice(L2, 0.124, 0.0, -1)
    lat2.load_gauge(dev(U2))
    other = LearnedPreconditioner(lat2, chan, seed=142)
    other.load(path)
    assert np.array_equal(other.get_weights(), W)
    x = orc.gauss((1, 2, orc.vol(L2) // 2, 4, 3), 143)
    ref, _ = orc.net_forward(x, W, chan, U2, L2, -1)
    e = rel(host(other.forward(dev(x))).ravel(), ref)
    log(f"weights file: {len(raw)} bytes, reloaded network vs oracle {e:.1e}")
    assert e < 2e-5
    wrong = LearnedPreconditioner(lat2, (1, 4, 1), seed=144)
    with pytest.raises(QcuError) as err:
        wrong.load(path)
    assert err.value.code == 1 and "channel" in str(err.value)
    (tmp_path / "short.qcunet").write_bytes(raw[:-8])
    with pytest.raises(QcuError) as err:
        other.load(tmp_path / "short.qcunet")
    assert err.value.code == 1
    assert np.array_equal(other.get_weights(), W)  # a refused file leaves the weights unchanged
    for n in (net, other, wrong):
        n.close()
