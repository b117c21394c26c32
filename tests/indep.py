"""Independent numpy formulation of the lattice operators, used to pin the oracle.

Written from the textbook definitions, not from the oracle's code: lexicographic
fields, full 3x3 links, projectors built as matrices from the DeGrand-Rossi
gamma matrices that PyQuda uses (python/PyQuda/pyquda/utils/gamma.py:19-64),
shifts by numpy.roll, and the checkerboard maps of
python/PyQuda/pyquda/field.py:22-63 restated with index arithmetic.
"""
import numpy as np

GAMMA = [
    np.array([[0, 0, 0, 1j], [0, 0, 1j, 0], [0, -1j, 0, 0], [-1j, 0, 0, 0]]),
    np.array([[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]], complex),
    np.array([[0, 0, 1j, 0], [0, 0, 0, -1j], [-1j, 0, 0, 0], [0, 1j, 0, 0]]),
    np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]], complex),
]
GAMMA5 = GAMMA[0] @ GAMMA[1] @ GAMMA[2] @ GAMMA[3]
ONE4 = np.eye(4, dtype=complex)

# numpy axis of each direction in a lexicographic [t][z][y][x] field
AX = {0: 3, 1: 2, 2: 1, 3: 0}


def to_lexico(cb, L):
    """[2][Lt][Lz][Ly][Lx/2][...] -> [Lt][Lz][Ly][Lx][...]"""
    Lx, Ly, Lz, Lt = L
    assert cb.shape[:5] == (2, Lt, Lz, Ly, Lx // 2)
    out = np.zeros((Lt, Lz, Ly, Lx) + cb.shape[5:], cb.dtype)
    t, z, y, x = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly), np.arange(Lx),
                             indexing="ij")
    par = (x + y + z + t) % 2
    out[...] = cb[par, t, z, y, x // 2]
    return out


def to_cb(lex, L):
    Lx, Ly, Lz, Lt = L
    out = np.zeros((2, Lt, Lz, Ly, Lx // 2) + lex.shape[4:], lex.dtype)
    t, z, y, x = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly), np.arange(Lx),
                             indexing="ij")
    par = (x + y + z + t) % 2
    out[par, t, z, y, x // 2] = lex
    return out


def gauge_to_lexico(U, L):
    """external gauge [4][2][Vh][3][3] -> [4][Lt][Lz][Ly][Lx][3][3]"""
    Lx, Ly, Lz, Lt = L
    U = U.reshape(4, 2, Lt, Lz, Ly, Lx // 2, 3, 3)
    return np.stack([to_lexico(U[mu], L) for mu in range(4)])


def fermion_to_lexico(psi, L):
    Lx, Ly, Lz, Lt = L
    return to_lexico(psi.reshape(2, Lt, Lz, Ly, Lx // 2, 4, 3), L)


def fermion_to_cb(psi, L):
    return to_cb(psi, L).reshape(2, -1, 4, 3)


def _phase(L, mu, fwd, t_boundary):
    """sign on hops that cross the time boundary"""
    Lt = L[3]
    ph = np.ones((Lt, 1, 1, 1))
    if mu == 3 and t_boundary == -1:
        ph[Lt - 1 if fwd else 0] = -1
    return ph


def hop_lexico(psi, Ul, L, t_boundary=1):
    """D psi(x) = sum_mu (1-g_mu) U_mu(x) psi(x+mu) + (1+g_mu) U_mu(x-mu)^+ psi(x-mu)"""
    out = np.zeros_like(psi)
    for mu in range(4):
        ax = AX[mu]
        fwd_psi = np.roll(psi, -1, axis=ax)
        t = np.einsum("tzyxab,tzyxsb->tzyxsa", Ul[mu], fwd_psi)
        t = np.einsum("rs,tzyxsa->tzyxra", ONE4 - GAMMA[mu], t)
        out += _phase(L, mu, True, t_boundary)[..., None, None] * t
        Ub = np.roll(Ul[mu], 1, axis=ax)
        bwd_psi = np.roll(psi, 1, axis=ax)
        t = np.einsum("tzyxba,tzyxsb->tzyxsa", Ub.conj(), bwd_psi)
        t = np.einsum("rs,tzyxsa->tzyxra", ONE4 + GAMMA[mu], t)
        out += _phase(L, mu, False, t_boundary)[..., None, None] * t
    return out


def sigma(mu, nu):
    return 0.5j * (GAMMA[mu] @ GAMMA[nu] - GAMMA[nu] @ GAMMA[mu])


def _mm(a, b):
    return np.einsum("...ab,...bc->...ac", a, b)


def _dag(a):
    return np.conj(np.swapaxes(a, -1, -2))


def clover_lexico(Ul, coeff):
    """A(x) = 1 - coeff sum_{mu<nu} sigma_munu F_munu, F = (Q - Q^+)/(8i), Q = 4 leaves.

    Returns [Lt][Lz][Ly][Lx][12][12] indexed (s*3+c, s'*3+c')."""
    shp = Ul.shape[1:5]
    A = np.zeros(shp + (4, 3, 4, 3), complex)

    def sh(f, mu, d):  # f(x + d*mu)
        return np.roll(f, -d, axis=AX[mu])

    for mu in range(4):
        for nu in range(mu + 1, 4):
            Um, Un = Ul[mu], Ul[nu]
            Q = _mm(_mm(_mm(Um, sh(Un, mu, 1)), _dag(sh(Um, nu, 1))), _dag(Un))
            Q = Q + _mm(_mm(_mm(Un, _dag(sh(sh(Um, mu, -1), nu, 1))), _dag(sh(Un, mu, -1))),
                        sh(Um, mu, -1))
            Q = Q + _mm(_mm(_mm(_dag(sh(Um, mu, -1)), _dag(sh(sh(Un, mu, -1), nu, -1))),
                            sh(sh(Um, mu, -1), nu, -1)), sh(Un, nu, -1))
            Q = Q + _mm(_mm(_mm(_dag(sh(Un, nu, -1)), sh(Um, nu, -1)), sh(sh(Un, mu, 1), nu, -1)),
                        _dag(Um))
            F = (Q - _dag(Q)) / 8j
            A -= coeff * np.einsum("rs,...ab->...rasb", sigma(mu, nu), F)
    A += np.einsum("rs,ab->rasb", ONE4, np.eye(3))
    return A.reshape(shp + (12, 12))


def clover_cb(Ul, L, coeff):
    """clover in the external per-parity layout [2][Vh][12][12]"""
    return to_cb(clover_lexico(Ul, coeff), L).reshape(2, -1, 12, 12)


def mat_lexico(psi, Ul, L, kappa, coeff=None, t_boundary=1):
    """M psi = A psi - kappa D psi  (A = 1 when coeff is None)"""
    d = hop_lexico(psi, Ul, L, t_boundary)
    if coeff is None:
        a = psi
    else:
        A = clover_lexico(Ul, coeff)
        a = np.einsum("...ij,...j->...i", A, psi.reshape(psi.shape[:4] + (12,))).reshape(psi.shape)
    return a - kappa * d


def dense_matrix(op, n):
    """dense matrix of a linear map on C^n given as a function of a flat vector"""
    M = np.zeros((n, n), complex)
    e = np.zeros(n, complex)
    for k in range(n):
        e[k] = 1
        M[:, k] = op(e)
        e[k] = 0
    return M


def random_su3_gauge_transform(L, rng):
    """g(x) in SU(3) from QR of Gaussian matrices, lexicographic [Lt][Lz][Ly][Lx][3][3]"""
    Lx, Ly, Lz, Lt = L
    z = rng.normal(size=(Lt, Lz, Ly, Lx, 3, 3)) + 1j * rng.normal(size=(Lt, Lz, Ly, Lx, 3, 3))
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=-2, axis2=-1)
    q = q * (d / np.abs(d))[..., None, :]
    det = np.linalg.det(q)
    q[..., :, 0] /= det[..., None]
    return q


def gauge_transform(Ul, g):
    out = np.empty_like(Ul)
    for mu in range(4):
        out[mu] = _mm(_mm(g, Ul[mu]), _dag(np.roll(g, -1, axis=AX[mu])))
    return out


def gauge_from_lexico(Ul, L):
    """[4][Lt][Lz][Ly][Lx][3][3] -> external [4][2][Vh][3][3]"""
    return np.ascontiguousarray(np.stack([to_cb(Ul[mu], L).reshape(2, -1, 3, 3)
                                          for mu in range(4)]))


def net_lexico(inp, Ws, Ul, L, t_boundary=1):
    """Gauge-covariant linear network on lexicographic channel fields.

    inp: [cin][Lt][Lz][Ly][Lx][4][3]; Ws: per layer [9][cout][cin][4][4] with
    term 2 mu = W U_mu(x) psi(x+mu), 2 mu + 1 = W U_mu(x-mu)^+ psi(x-mu), 8 = W psi(x).
    """
    a = inp
    for W in Ws:
        terms = []
        for mu in range(4):
            ax = AX[mu]
            f = np.einsum("tzyxab,ctzyxsb->ctzyxsa", Ul[mu], np.roll(a, -1, axis=ax + 1))
            terms.append(_phase(L, mu, True, t_boundary)[None, ..., None, None] * f)
            Ub = np.roll(Ul[mu], 1, axis=ax)
            b = np.einsum("tzyxba,ctzyxsb->ctzyxsa", Ub.conj(), np.roll(a, 1, axis=ax + 1))
            terms.append(_phase(L, mu, False, t_boundary)[None, ..., None, None] * b)
        terms.append(a)
        a = sum(np.einsum("oirs,itzyxsc->otzyxrc", W[t], terms[t]) for t in range(9))
    return a
