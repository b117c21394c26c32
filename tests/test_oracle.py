"""Pins the FP64 host oracle against mathematics that fixes each routine.

No reference golden vectors exist for this path (the reference's tests compare
against QUDA at run time and store no data); every check below is an
independent formulation, a closed form or an invariant chosen so that a
dropped term, a wrong sign or a transposed operand fails.
"""
import numpy as np
import pytest
import scipy.linalg

import indep

L44 = (4, 4, 4, 4)
L448 = (4, 4, 4, 8)
L2 = (2, 2, 2, 2)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# ---------------------------------------------------------------- generators
def test_su3_exp_matches_expm(orc):
    rng = np.random.default_rng(1)
    lam = np.zeros((8, 3, 3), complex)
    lam[0][0, 1] = lam[0][1, 0] = 1
    lam[1][0, 1], lam[1][1, 0] = -1j, 1j
    lam[2][0, 0], lam[2][1, 1] = 1, -1
    lam[3][0, 2] = lam[3][2, 0] = 1
    lam[4][0, 2], lam[4][2, 0] = -1j, 1j
    lam[5][1, 2] = lam[5][2, 1] = 1
    lam[6][1, 2], lam[6][2, 1] = -1j, 1j
    lam[7] = np.diag([1, 1, -2]) / np.sqrt(3)
    for eps in (0.05, 0.3, 1.7):
        w = rng.normal(size=8)
        H = np.einsum("a,aij->ij", w, lam)
        ref = scipy.linalg.expm(1j * eps * H)
        assert np.abs(orc.su3_exp(w, eps) - ref).max() < 1e-14


@pytest.mark.parametrize("eps", [0.2, 0.3])
def test_gauge_random_is_su3(orc, eps):
    U = orc.gauge_random(L448, eps, 1234)
    m = U.reshape(-1, 3, 3)
    uu = np.einsum("nba,nbc->nac", m.conj(), m)
    assert np.abs(uu - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(m) - 1).max() < 1e-14
    # reproducible and seed dependent
    assert np.array_equal(U, orc.gauge_random(L448, eps, 1234))
    assert not np.allclose(U, orc.gauge_random(L448, eps, 1235))
    # fluctuation scales with eps: |1 - Re tr U / 3| ~ eps^2
    d = 1 - np.einsum("nii->n", m).real.mean() / 3
    assert 0.1 * eps**2 < d < 5 * eps**2


def test_gauss_moments(orc):
    v = orc.gauss(200000, 7)
    assert abs(v.real.mean()) < 0.01 and abs(v.imag.mean()) < 0.01
    assert abs(v.real.var() - 1) < 0.02 and abs(v.imag.var() - 1) < 0.02
    assert abs(np.mean(v.real * v.imag)) < 0.01


# ---------------------------------------------------------------- hop
@pytest.mark.parametrize("L", [L2, L44, L448, (2, 4, 6, 4)])
@pytest.mark.parametrize("tb", [1, -1])
def test_hop_matches_independent_lexicographic(orc, L, tb):
    U = orc.gauge_random(L, 0.3, 11)
    psi = orc.gauss((2, orc.vol(L) // 2, 4, 3), 12)
    out = np.empty_like(psi)
    out[0] = orc.hop(np.ascontiguousarray(psi[1]), U, L, 0, tb)
    out[1] = orc.hop(np.ascontiguousarray(psi[0]), U, L, 1, tb)
    Ul = indep.gauge_to_lexico(U, L)
    ref = indep.fermion_to_cb(indep.hop_lexico(indep.fermion_to_lexico(psi, L), Ul, L, tb), L)
    assert rel(out, ref) < 1e-14
    assert rel(orc.hop_full(psi, U, L, tb), ref) < 1e-14


def test_hop_free_field_plane_wave(orc):
    """U = 1: D e^{ipx} u = sum_mu [2 cos p_mu - 2i g_mu sin p_mu] e^{ipx} u"""
    L = (4, 6, 4, 8)
    Lx, Ly, Lz, Lt = L
    U = np.zeros((4, 2, orc.vol(L) // 2, 3, 3), complex)
    U[...] = np.eye(3)
    rng = np.random.default_rng(3)
    for tb in (1, -1):
        n = rng.integers(0, 4, size=4)
        p = [2 * np.pi * n[0] / Lx, 2 * np.pi * n[1] / Ly, 2 * np.pi * n[2] / Lz,
             (2 * n[3] + (1 if tb == -1 else 0)) * np.pi / Lt]
        u = rng.normal(size=(4, 3)) + 1j * rng.normal(size=(4, 3))
        t, z, y, x = np.meshgrid(*[np.arange(v) for v in (Lt, Lz, Ly, Lx)], indexing="ij")
        ph = np.exp(1j * (p[0] * x + p[1] * y + p[2] * z + p[3] * t))
        psi = indep.fermion_to_cb(ph[..., None, None] * u, L)
        K = sum(2 * np.cos(p[m]) * indep.ONE4 - 2j * np.sin(p[m]) * indep.GAMMA[m] for m in range(4))
        ref = indep.fermion_to_cb(ph[..., None, None] * (K @ u), L)
        assert rel(orc.hop_full(psi, U, L, tb), ref) < 1e-13


def test_hop_gauge_covariance(orc):
    L = L448
    rng = np.random.default_rng(5)
    U = orc.gauge_random(L, 0.3, 21)
    Ul = indep.gauge_to_lexico(U, L)
    g = indep.random_su3_gauge_transform(L, rng)
    Ug = indep.gauge_from_lexico(indep.gauge_transform(Ul, g), L)
    psi = orc.gauss((2, orc.vol(L) // 2, 4, 3), 22)
    pl = indep.fermion_to_lexico(psi, L)
    psig = indep.fermion_to_cb(np.einsum("tzyxab,tzyxsb->tzyxsa", g, pl), L)
    lhs = orc.hop_full(psig, Ug, L, -1)
    rhs = indep.fermion_to_cb(np.einsum("tzyxab,tzyxsb->tzyxsa", g,
                                        indep.fermion_to_lexico(orc.hop_full(psi, U, L, -1), L)), L)
    assert rel(lhs, rhs) < 1e-13


def test_hop_gamma5_hermiticity(orc):
    L = L448
    U = orc.gauge_random(L, 0.3, 31)
    vh = orc.vol(L) // 2
    chi = orc.gauss((vh, 4, 3), 32)
    psi = orc.gauss((vh, 4, 3), 33)
    g5 = np.diag(indep.GAMMA5)  # diagonal in DeGrand-Rossi: (1, 1, -1, -1)
    assert np.allclose(indep.GAMMA5, np.diag(g5))
    lhs = np.vdot(chi, orc.hop(psi, U, L, 0, -1))
    rhs = np.vdot(g5[None, :, None] * orc.hop(g5[None, :, None] * chi, U, L, 1, -1), psi)
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


# ---------------------------------------------------------------- clover
@pytest.mark.parametrize("coeff", [1.0, 0.124])
def test_clover_matches_textbook(orc, coeff):
    L = L448
    U = orc.gauge_random(L, 0.3, 41)
    Ul = indep.gauge_to_lexico(U, L)
    ref = indep.clover_cb(Ul, L, coeff)
    for p in (0, 1):
        A = orc.clover_build(U, L, p, coeff)
        assert rel(A, ref[p]) < 1e-14


def test_clover_unit_gauge_and_structure(orc):
    L = L44
    U = np.zeros((4, 2, orc.vol(L) // 2, 3, 3), complex)
    U[...] = np.eye(3)
    A = orc.clover_build(U, L, 0, 1.0)
    assert np.abs(A - np.eye(12)).max() == 0
    U = orc.gauge_random(L, 0.3, 42)
    A = orc.clover_build(U, L, 1, 1.0)
    assert np.abs(A - np.conj(np.swapaxes(A, 1, 2))).max() < 1e-14  # Hermitian
    blk = A.reshape(-1, 4, 3, 4, 3)
    assert np.abs(blk[:, 0:2, :, 2:4, :]).max() < 1e-15  # chiral blocks {0,1},{2,3}
    assert np.abs(blk[:, 2:4, :, 0:2, :]).max() < 1e-15


def test_clover_gauge_covariance(orc):
    L = L448
    rng = np.random.default_rng(6)
    U = orc.gauge_random(L, 0.3, 43)
    Ul = indep.gauge_to_lexico(U, L)
    g = indep.random_su3_gauge_transform(L, rng)
    Ug = indep.gauge_from_lexico(indep.gauge_transform(Ul, g), L)
    gcb = indep.to_cb(g, L).reshape(2, -1, 3, 3)
    for p in (0, 1):
        A = orc.clover_build(U, L, p, 1.0).reshape(-1, 4, 3, 4, 3)
        Ag = orc.clover_build(Ug, L, p, 1.0).reshape(-1, 4, 3, 4, 3)
        rot = np.einsum("nab,nrbsc,ndc->nrasd", gcb[p], A, gcb[p].conj())
        assert np.abs(Ag - rot).max() < 1e-13


def test_clover_invert(orc):
    L = L44
    U = orc.gauge_random(L, 0.3, 44)
    A = orc.clover_build(U, L, 0, 1.0)
    Ai = orc.clover_invert(A)
    cond = np.linalg.cond(A).max()
    assert np.abs(np.einsum("nij,njk->nik", A, Ai) - np.eye(12)).max() < 1e-15 * cond * 12
    assert rel(Ai, np.linalg.inv(A)) < 1e-15 * cond * 12


def test_clover_dslash_composition(orc):
    L = L448
    U = orc.gauge_random(L, 0.3, 45)
    Ul = indep.gauge_to_lexico(U, L)
    psi = orc.gauss((2, orc.vol(L) // 2, 4, 3), 46)
    d = indep.fermion_to_cb(indep.hop_lexico(indep.fermion_to_lexico(psi, L), Ul, L, 1), L)
    A = indep.clover_cb(Ul, L, 1.0)
    for p in (0, 1):
        ref = np.linalg.solve(A[p], d[p].reshape(-1, 12, 1)).reshape(-1, 4, 3)
        out = orc.clover_dslash(np.ascontiguousarray(psi[1 - p]), U, L, p, 1.0, 1)
        # kappa*csw = 1 at eps = 0.3 gives cond(A) ~ 1e2; unpivoted Gauss-Jordan
        assert rel(out, ref) < 1e-12


# ---------------------------------------------------------------- solver
def _dense_full(L, U, kappa, csw, tb):
    Ul = indep.gauge_to_lexico(U, L)
    n = orc_vol = L[0] * L[1] * L[2] * L[3] * 12
    coeff = None if csw == 0 else kappa * csw

    def op(v):
        psi = indep.fermion_to_lexico(v.reshape(2, -1, 4, 3), L)
        return indep.fermion_to_cb(indep.mat_lexico(psi, Ul, L, kappa, coeff, tb), L).reshape(-1)
    return indep.dense_matrix(op, n)


@pytest.mark.parametrize("csw", [0.0, 1.0])
def test_mat_full_matches_dense(orc, csw):
    L = L2
    U = orc.gauge_random(L, 0.3, 51)
    M = _dense_full(L, U, 0.124, csw, -1)
    x = orc.gauss((2, orc.vol(L) // 2, 4, 3), 52)
    assert rel(orc.mat_full(x, U, L, 0.124, csw, -1).reshape(-1), M @ x.reshape(-1)) < 1e-14


@pytest.mark.parametrize("L,csw,tb", [(L2, 0.0, 1), (L2, 1.0, -1), (L44, 1.0, -1), (L44, 0.0, -1)])
def test_bicgstab_matches_dense_solve(orc, L, csw, tb):
    kappa = 0.124
    U = orc.gauge_random(L, 0.3, 61)
    b = orc.gauss((2, orc.vol(L) // 2, 4, 3), 62)
    x, info = orc.bicgstab(b, U, L, kappa, csw, tb, tol=1e-13, max_iter=2000)
    M = _dense_full(L, U, kappa, csw, tb)
    xd = np.linalg.solve(M, b.reshape(-1))
    assert rel(x.reshape(-1), xd) < 1e-11
    assert info["true_res"] < 1e-12
    assert rel(M @ x.reshape(-1), b.reshape(-1)) < 1e-12


def test_bicgstab_reference_stopping_rule(orc):
    """mode 0: stop on Re<r,r> < tol absolute, as include/lattice_bistabcg.h:301"""
    L = L44
    U = orc.gauge_random(L, 0.3, 63)
    b = orc.gauss((2, orc.vol(L) // 2, 4, 3), 64)
    x, info = orc.bicgstab(b, U, L, 0.125, 0.0, 1, tol=1e-9, max_iter=1000, mode=0)
    bh_norm2_est = np.vdot(b, b).real  # |b_hat| is of the same order
    assert info["schur_res"] ** 2 * bh_norm2_est < 1e-7
    assert info["iters"] < 1000


def test_blas_closed_forms(orc):
    """<a, b> = sum conj(a) b and |a|^2 against numpy on an odd length"""
    a = orc.gauss((12345,), 1)
    b = orc.gauss((12345,), 2)
    assert abs(orc.dot(a, b) - np.vdot(a, b)) < 1e-12 * abs(np.vdot(a, b))
    assert abs(orc.norm2(a) - np.vdot(a, a).real) < 1e-12 * np.vdot(a, a).real


def test_true_residual_closed_form(orc):
    """|b - M x| / |b| from the dense operator of the independent restatement"""
    L, kappa, csw, tb = (4, 4, 4, 4), 0.124, 1.0, -1
    U = orc.gauge_random(L, 0.3, 3)
    x = orc.gauss((2, orc.vol(L) // 2, 4, 3), 4)
    b = orc.gauss((2, orc.vol(L) // 2, 4, 3), 5)
    Ul = indep.gauge_to_lexico(U, L)
    mx = indep.fermion_to_cb(indep.mat_lexico(indep.fermion_to_lexico(x, L), Ul, L, kappa, kappa * csw, tb), L)
    want = np.linalg.norm(b - mx) / np.linalg.norm(b)
    got = orc.true_residual(x, b, U, L, kappa, csw, tb)
    assert abs(got - want) < 1e-12 * want
