"""Handle-based host interface over libqcu.so for torch device tensors.

Fields use the external (PyQuda) layouts of qcu's C ABI:
  gauge   complex128 [4][2][Lt][Lz][Ly][Lx/2][3][3]
  half    complex128 [Lt][Lz][Ly][Lx/2][4][3]   (one parity)
  full    complex128 [2][Lt][Lz][Ly][Lx/2][4][3]
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import SolveInfo, SolveParam

FP64, FP32, FP16 = 0, 1, 2
PREC = {"fp64": FP64, "fp32": FP32, "fp16": FP16}
METHODS = {"bicgstab": 0, "cgnr": 1}


def _ptr(t: torch.Tensor) -> C.c_void_p:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError("expected a GPU torch tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def _check_field(t, n_complex: int, what: str):
    """A field handed to the library: a contiguous complex128 tensor of exactly
    n_complex elements (the library reads and writes that many)."""
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{what}: expected a torch tensor, got {type(t).__name__}")
    if t.dtype != torch.complex128:
        raise ValueError(f"{what}: expected complex128, got {t.dtype}")
    if t.numel() != n_complex:
        raise ValueError(f"{what}: expected {n_complex} complex values, got {t.numel()}")
    if not t.is_contiguous():
        raise ValueError(f"{what}: expected a contiguous tensor")


def _field_ptr(t, n_complex: int, what: str) -> C.c_void_p:
    _check_field(t, n_complex, what)
    return _ptr(t)


def vol(L):
    return int(L[0] * L[1] * L[2] * L[3])


def gauge_shape(L):
    Lx, Ly, Lz, Lt = L
    return (4, 2, Lt, Lz, Ly, Lx // 2, 3, 3)


def fermion_shape(L, full=True):
    Lx, Ly, Lz, Lt = L
    return ((2,) if full else ()) + (Lt, Lz, Ly, Lx // 2, 4, 3)


def random_gauge(L, eps, seed, device="cuda"):
    """U = exp(i eps H), H = sum_a w_a lambda_a with w_a ~ N(0,1), generated on the GPU."""
    U = torch.empty(gauge_shape(L), dtype=torch.complex128, device=device)
    _lib.check(_lib.load().qcuGaugeRandom(_ptr(U), (C.c_int * 4)(*L), float(eps), int(seed)))
    return U


def gaussian(shape, seed, device="cuda"):
    """complex128 field with real and imaginary parts ~ N(0,1), generated on the GPU."""
    v = torch.empty(shape, dtype=torch.complex128, device=device)
    _lib.check(_lib.load().qcuGaussian(_ptr(v), v.numel(), int(seed)))
    return v


def point_source(L, spin=0, colour=0, device="cuda"):
    """1 at site (0,0,0,0), as python/PyQuda/pyquda/utils/source.py:7-22"""
    b = torch.zeros(fermion_shape(L), dtype=torch.complex128, device=device)
    b[0, 0, 0, 0, 0, spin, colour] = 1
    return b


@dataclass
class SolveResult:
    iters: int
    reliable_updates: int
    converged: bool
    fallback: bool
    true_res: float
    seconds: float
    restarts: int


class Lattice:
    """Resident gauge field (+ clover) for one (L, kappa, csw, t_boundary)."""

    def __init__(self, L, kappa, csw=0.0, t_boundary=-1):
        self.L = tuple(int(v) for v in L)
        self.kappa = float(kappa)
        self.csw = float(csw)
        self.t_boundary = int(t_boundary)
        self._lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self._lib.qcuLatticeCreate(C.byref(h), (C.c_int * 4)(*self.L), self.kappa,
                                              self.csw, self.t_boundary))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self._lib.qcuLatticeDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def vh(self):
        return vol(self.L) // 2

    @property
    def n_full(self):
        """complex values of a full-lattice spinor field"""
        return 12 * vol(self.L)

    @property
    def n_half(self):
        """complex values of a one-parity spinor field"""
        return 6 * vol(self.L)

    def load_gauge(self, U: torch.Tensor):
        pu = _field_ptr(U, 36 * vol(self.L), "gauge")
        _lib.check(self._lib.qcuLatticeLoadGauge(self.h, pu))
        torch.cuda.synchronize()

    def dslash(self, inp: torch.Tensor, parity: int, clover=False, prec=FP64) -> torch.Tensor:
        pin = _field_ptr(inp, self.n_half, "dslash input")
        out = torch.empty_like(inp)
        _lib.check(self._lib.qcuDslash(self.h, _ptr(out), pin, int(parity), int(clover),
                                       int(prec)))
        return out

    def mat(self, x: torch.Tensor) -> torch.Tensor:
        px = _field_ptr(x, self.n_full, "mat input")
        out = torch.empty_like(x)
        _lib.check(self._lib.qcuMatFull(self.h, _ptr(out), px))
        return out

    def true_residual(self, x, b) -> float:
        px, pb = _field_ptr(x, self.n_full, "x"), _field_ptr(b, self.n_full, "b")
        r = C.c_double()
        _lib.check(self._lib.qcuTrueResidual(self.h, px, pb, C.byref(r)))
        return r.value

    def solve(self, b, tol=1e-10, max_iter=20000, prec=FP64, reliable_delta=0.1, chunk=0,
              use_graph=True, precond=None, legacy_stop=False, x=None, method="bicgstab"):
        """x with |b - M x| / |b| <= tol; method "bicgstab" (any precision,
        optional preconditioner) or "cgnr" (CG on the normal equations, FP64)"""
        pb = _field_ptr(b, self.n_full, "b")
        if x is None:
            x = torch.empty_like(b)
        px = _field_ptr(x, self.n_full, "x")
        p = SolveParam(tol=tol, max_iter=max_iter, precision=int(prec),
                       reliable_delta=reliable_delta, chunk=chunk, use_graph=int(use_graph),
                       precond=None if precond is None else precond.handle,
                       legacy_stop=int(legacy_stop), method=METHODS[method])
        info = SolveInfo()
        _lib.check(self._lib.qcuSolve(self.h, px, pb, C.byref(p), C.byref(info)))
        return x, SolveResult(info.iters, info.reliable_updates, bool(info.converged),
                              bool(info.fallback), info.true_res, info.seconds, info.restarts)

    def bench_hop(self, prec=FP64, clover=None, xpay=True, iters=50, block=0, remap=-1):
        """average ms per launch of the resident hop kernel, timed with HIP events"""
        if clover is None:
            clover = self.csw != 0.0
        ms = C.c_double()
        _lib.check(self._lib.qcuBenchHop(self.h, int(prec), int(bool(clover)), int(xpay),
                                         int(iters), int(block), int(remap), C.byref(ms)))
        return ms.value

    def tune(self, prec=FP64):
        ms = C.c_double()
        _lib.check(self._lib.qcuTune(self.h, int(prec), C.byref(ms)))
        return dict(self.launch_shape(prec), ms=ms.value)

    def launch_shape(self, prec=FP64):
        """the hop kernel's launch shape used for this lattice and precision"""
        b, r, g = C.c_int(), C.c_int(), C.c_int()
        _lib.check(self._lib.qcuTuneGet(self.h, int(prec), C.byref(b), C.byref(r), C.byref(g)))
        return {"block": b.value, "remap": r.value, "group": g.value}

    def set_launch_shape(self, prec, block, remap, group):
        """pin the hop kernel's launch shape (as the tuner would choose it)"""
        _lib.check(self._lib.qcuTuneSet(self.h, int(prec), int(block), int(remap), int(group)))


class Batch:
    """Many systems solved together: nrhs[k] right-hand sides on lattices[k].

    Every kernel launch covers all systems (one launch per kernel for the
    whole batch; the right-hand sides of one lattice share link loads).  The
    lattices must stay alive and keep their gauge fields while the batch is
    used.  System order: lattice 0 rhs 0.., lattice 1 rhs 0.., ...
    """

    def __init__(self, lattices, nrhs):
        self._lib = _lib.load()
        self.lattices = list(lattices)
        self.nrhs = [int(n) for n in nrhs]
        hs = (C.c_void_p * len(self.lattices))(*[l.h.value for l in self.lattices])
        nr = (C.c_int * len(self.nrhs))(*self.nrhs)
        h = C.c_void_p()
        _lib.check(self._lib.qcuBatchCreate(C.byref(h), hs, nr, len(self.lattices)))
        self.h = h
        n = C.c_int()
        _lib.check(self._lib.qcuBatchSystems(self.h, C.byref(n)))
        self.nsys = n.value

    def close(self):
        if getattr(self, "h", None):
            self._lib.qcuBatchDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, bs, tol=1e-10, max_iter=20000, prec=FP64, reliable_delta=0.1, chunk=0,
              use_graph=True, xs=None, method="bicgstab"):
        if len(bs) != self.nsys or (xs is not None and len(xs) != self.nsys):
            raise ValueError(f"expected {self.nsys} right-hand sides (and solutions)")
        sizes = [lat.n_full for lat, n in zip(self.lattices, self.nrhs) for _ in range(n)]
        bp = (C.c_void_p * self.nsys)(*[_field_ptr(b, n, "b").value for b, n in zip(bs, sizes)])
        if xs is None:
            xs = [torch.empty_like(b) for b in bs]
        xp = (C.c_void_p * self.nsys)(*[_field_ptr(x, n, "x").value for x, n in zip(xs, sizes)])
        p = SolveParam(tol=tol, max_iter=max_iter, precision=int(prec), reliable_delta=reliable_delta,
                       chunk=chunk, use_graph=int(use_graph), precond=None, legacy_stop=0,
                       method=METHODS[method])
        info = (SolveInfo * self.nsys)()
        _lib.check(self._lib.qcuBatchSolve(self.h, xp, bp, C.byref(p), info))
        res = [SolveResult(i.iters, i.reliable_updates, bool(i.converged), bool(i.fallback), i.true_res,
                           i.seconds, i.restarts) for i in info]
        return xs, res

    def bench_hop(self, prec=FP64, xpay=True, iters=50):
        ms = C.c_double()
        _lib.check(self._lib.qcuBatchBenchHop(self.h, int(prec), int(xpay), int(iters), C.byref(ms)))
        return ms.value

    def tune(self, prec=FP64):
        ms = C.c_double()
        _lib.check(self._lib.qcuBatchTune(self.h, int(prec), C.byref(ms)))
        return ms.value

    def set_launch_shape(self, prec, block, remap, group):
        """pin the batched hop kernel's launch shape"""
        _lib.check(self._lib.qcuBatchTuneSet(self.h, int(prec), int(block), int(remap), int(group)))


def tune_save(path):
    _lib.check(_lib.load().qcuTuneSave(str(path).encode()))


def tune_load(path):
    _lib.check(_lib.load().qcuTuneLoad(str(path).encode()))


def schur_solutions(lattice: Lattice, sources, tol=1e-8, prec=FP16, reliable_delta=0.01):
    """y = Mhat^-1 c for odd-site sources c ([Lt][Lz][Ly][Lx/2][4][3] or [Vh][4][3]):
    the full system M x = (0, A_o c) has the Schur right-hand side
    bhat = A_o^-1 (A_o c) = c, so its odd half solves Mhat y = c.  A_o c is the
    odd half of M applied to (0, c).  Returns the list of y (complex128, shape of c)."""
    full = fermion_shape(lattice.L)
    out = []
    for c in sources:
        z = torch.zeros(full, dtype=torch.complex128, device=c.device)
        z[1] = c.reshape(full[1:])
        bf = torch.zeros_like(z)
        bf[1] = lattice.mat(z)[1]
        x, res = lattice.solve(bf, tol=tol, prec=prec, reliable_delta=reliable_delta)
        if not res.converged:
            raise RuntimeError(f"schur_solutions: solve did not reach tol {tol}: {res}")
        out.append(x[1].reshape(c.shape).clone())
    return out


class LearnedPreconditioner:
    """Gauge-covariant linear network K on one lattice (see qcuNetCreate).

    channels: (1, 8, 8, 8, 1) gives 4 layers with 8 hidden channels.  Weights
    are complex, in the order layer, term (2 mu forward hop, 2 mu + 1 backward
    hop, 8 site), output channel, input channel, 4x4 spin matrix.  Pass the
    object as ``precond=`` to :meth:`Lattice.solve`.
    """

    def __init__(self, lattice: Lattice, channels=(1, 8, 8, 8, 1), seed=1, init_scale=0.05):
        self._lib = _lib.load()
        self.lattice = lattice
        self.channels = tuple(int(c) for c in channels)
        ch = (C.c_int * len(self.channels))(*self.channels)
        h = C.c_void_p()
        _lib.check(self._lib.qcuNetCreate(C.byref(h), lattice.h, len(self.channels) - 1, ch,
                                          int(seed), float(init_scale)))
        self.handle = h
        self.nweights = int(self._lib.qcuNetWeightCount(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            self._lib.qcuNetDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _host_buf(self):
        import numpy as np
        return np.zeros(self.nweights, dtype=np.complex128)

    def get_weights(self):
        w = self._host_buf()
        _lib.check(self._lib.qcuNetGetWeights(self.handle, w.ctypes.data_as(C.POINTER(C.c_double)),
                                              self.nweights))
        return w

    def set_weights(self, w):
        import numpy as np
        w = np.ascontiguousarray(np.asarray(w, dtype=np.complex128).reshape(-1))
        if w.size != self.nweights:
            raise ValueError(f"expected {self.nweights} weights, got {w.size}")
        _lib.check(self._lib.qcuNetSetWeights(self.handle, w.ctypes.data_as(C.POINTER(C.c_double)),
                                              self.nweights))

    def gradient(self):
        g = self._host_buf()
        _lib.check(self._lib.qcuNetGetGradient(self.handle, g.ctypes.data_as(C.POINTER(C.c_double)),
                                               self.nweights))
        return g

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """K x on a full-lattice field"""
        # the library requires one input and one output channel
        px = _field_ptr(x, self.lattice.n_full, "network input")
        out = torch.empty_like(x)
        _lib.check(self._lib.qcuNetForward(self.handle, _ptr(out), px))
        return out

    def backward(self, x: torch.Tensor, grad_out: torch.Tensor, want_input_grad=True):
        """(weight gradient as complex128 numpy, input gradient or None) for a
        real loss with dL/dRe(Kx) + i dL/dIm(Kx) = grad_out"""
        px = _field_ptr(x, self.lattice.n_full, "network input")
        pg = _field_ptr(grad_out, self.lattice.n_full, "output gradient")
        gi = torch.empty_like(x) if want_input_grad else None
        _lib.check(self._lib.qcuNetBackward(self.handle, None if gi is None else _ptr(gi), pg, px))
        return self.gradient(), gi

    def apply_odd(self, x_odd: torch.Tensor) -> torch.Tensor:
        """K on an odd-site field (even sites taken as zero), as in the solver"""
        px = _field_ptr(x_odd, self.lattice.n_half, "odd input")
        out = torch.empty_like(x_odd)
        _lib.check(self._lib.qcuNetApplyOdd(self.handle, _ptr(out), px))
        return out

    def save(self, path):
        """write the weights with their metadata (see qcuNetSave)"""
        _lib.check(self._lib.qcuNetSave(self.handle, str(path).encode()))

    def load(self, path):
        """read weights written by save() (same channel list)"""
        _lib.check(self._lib.qcuNetLoad(self.handle, str(path).encode()))

    def train(self, sources, epochs=10, batch=16, lr=1e-3):
        """Adam on |Mhat K b - b|^2 / |b|^2 over odd-site sources; returns
        (per-epoch mean losses, seconds)"""
        import numpy as np
        n = len(sources)
        sp = (C.c_void_p * n)(*[_field_ptr(s, self.lattice.n_half, "source").value for s in sources])
        losses = np.zeros(epochs)
        secs = C.c_double()
        _lib.check(self._lib.qcuNetTrain(self.handle, sp, n, int(epochs), int(batch), float(lr),
                                         losses.ctypes.data_as(C.POINTER(C.c_double)), C.byref(secs)))
        return losses, secs.value

    def train_pairs(self, sources, solutions, epochs=1, batch=4, lr=2e-3, solution_weight=1.0):
        """Adam on (source, solution) pairs (y = Mhat^-1 c, e.g. from schur_solutions):
        per pair the loss w |K c - y|^2 / |y|^2 + (1 - w) |Mhat K c - c|^2 / |c|^2,
        w = solution_weight.  The gradient of the first term comes from the
        library's forward and backward passes (qcuNetForward / qcuNetBackward),
        of the second from qcuNetLossGrad; Adam (beta 0.9, 0.999, eps 1e-8, on
        the batch mean, as qcuNetTrain) runs on the host.  Returns
        (per-epoch mean losses, seconds)."""
        import time
        import numpy as np
        if len(sources) != len(solutions) or not sources:
            raise ValueError("one solution per source")
        for c, y in zip(sources, solutions):
            _check_field(c, self.lattice.n_half, "source")
            _check_field(y, self.lattice.n_half, "solution")
        w = float(solution_weight)
        n, vh = len(sources), self.lattice.vh
        full = (1,) + fermion_shape(self.lattice.L)
        xin = torch.zeros(full, dtype=torch.complex128, device="cuda")
        gout = torch.zeros_like(xin)
        W = self.get_weights()
        m = np.zeros(2 * W.size)
        v = np.zeros(2 * W.size)
        step = 0
        losses = np.zeros(epochs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for e in range(epochs):
            lsum = 0.0
            for b0 in range(0, n, batch):
                b1 = min(n, b0 + batch)
                g = np.zeros(W.size, complex)
                for k in range(b0, b1):
                    c, y = sources[k], solutions[k]
                    if w > 0:
                        xin.zero_()
                        xin[0, 1] = c.view(full[2:])
                        err = self.forward(xin)[0, 1].reshape(vh, 4, 3) - y
                        y2 = float(torch.vdot(y.flatten(), y.flatten()).real)
                        lsum += w * float(torch.vdot(err.flatten(), err.flatten()).real) / y2
                        gout.zero_()
                        gout[0, 1] = (2.0 / y2) * err.view(full[2:])
                        gs, _ = self.backward(xin, gout, want_input_grad=False)
                        g += w * gs
                    if w < 1:
                        lr_, gr = self.loss_grad(c)
                        lsum += (1 - w) * lr_
                        g += (1 - w) * gr
                g /= b1 - b0
                step += 1
                gv = np.concatenate([g.real, g.imag])
                m = 0.9 * m + 0.1 * gv
                v = 0.999 * v + 0.001 * gv * gv
                d = lr * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
                W = W - (d[:W.size] + 1j * d[W.size:])
                self.set_weights(W)
            losses[e] = lsum / n
        torch.cuda.synchronize()
        return losses, time.perf_counter() - t0

    def loss_grad(self, source):
        """(loss, weight gradient) of |Mhat K b - b|^2 / |b|^2 for one odd source"""
        ps = _field_ptr(source, self.lattice.n_half, "source")
        out = C.c_double()
        _lib.check(self._lib.qcuNetLossGrad(self.handle, ps, C.byref(out)))
        return out.value, self.gradient()

    def bench_layer(self, layer, iters=10):
        """(ms per launch, algorithmic flops per launch) of one layer's forward kernel"""
        ms, fl = C.c_double(), C.c_double()
        _lib.check(self._lib.qcuNetBenchLayer(self.handle, int(layer), int(iters), C.byref(ms), C.byref(fl)))
        return ms.value, fl.value

    def bench_apply(self, iters=10):
        """ms per preconditioner application as the solver runs it (HIP events)"""
        ms = C.c_double()
        _lib.check(self._lib.qcuNetBenchApply(self.handle, int(iters), C.byref(ms)))
        return ms.value

    def bench_precond_layer(self, layer, iters=10):
        """ms per launch of one layer's kernels as the preconditioner runs them
        (HIP events)"""
        ms = C.c_double()
        _lib.check(self._lib.qcuNetBenchPrecondLayer(self.handle, int(layer), int(iters), C.byref(ms)))
        return ms.value

    def loss(self, sources) -> float:
        n = len(sources)
        sp = (C.c_void_p * n)(*[_field_ptr(s, self.lattice.n_half, "source").value for s in sources])
        out = C.c_double()
        _lib.check(self._lib.qcuNetLoss(self.handle, sp, n, C.byref(out)))
        return out.value
