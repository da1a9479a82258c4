"""Pins of the DF oracle (oracle/dferi.py, a restatement of AtomicIntegrals.jl:31-41, DFERI.jl:15-69
and Fock.jl:81-98): the identity-coefficient special case, four-index consistency on an AO tensor
that is not symmetric in (mu, nu), orthogonal invariance of the Frobenius norm, and for the DF fit
the four-index tensor against a linear solve, the unit metric and the squared metric."""
import numpy as np

from oracle.dferi import df_ao_to_mo, df_fit, rhf_eps


def random_ao(naux, nbf, seed, symmetric=True):
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((naux, nbf, nbf))
    if symmetric:
        b = b + np.transpose(b, (0, 2, 1))
    C, _ = np.linalg.qr(rng.standard_normal((nbf, nbf)))
    eps = np.sort(rng.uniform(-2.0, 3.0, nbf))
    return b, C, eps


def test_identity_coefficients_slice_the_ao_tensor():
    b, _, eps = random_ao(5, 9, 0, symmetric=False)
    nd, core, inac = 4, 1, 2
    Boo, Bov, Bvv = df_ao_to_mo(b, np.eye(9), nd, core, inac)
    assert np.array_equal(Boo, b[:, 1:4, 1:4])
    assert np.array_equal(Bov, b[:, 1:4, 4:7])
    assert np.array_equal(Bvv, b[:, 4:7, 4:7])
    eo, ev = rhf_eps(eps, nd, core, inac)
    assert np.array_equal(eo, eps[1:4]) and np.array_equal(ev, eps[4:7])


def test_four_index_consistency_nonsymmetric():
    b, C, _ = random_ao(6, 7, 1, symmetric=False)
    C = C + 0.1 * np.random.default_rng(9).standard_normal(C.shape)   # general (non-orthogonal) C
    nd, core, inac = 3, 1, 1
    Boo, Bov, Bvv = df_ao_to_mo(b, C, nd, core, inac)
    ao = np.einsum("Qmn,Qls->mnls", b, b)                               # (mu nu|la si)
    Co, Cv = C[:, 1:3], C[:, 3:6]
    x = np.tensordot(Co, ao, axes=(0, 0))                               # i nu la si
    x = np.tensordot(Cv, x, axes=(0, 1))                                # a i la si
    x = np.tensordot(Co, x, axes=(0, 2))                                # j a i si
    x = np.tensordot(Cv, x, axes=(0, 3))                                # b j a i
    iajb = np.transpose(x, (3, 2, 1, 0))
    assert np.max(np.abs(np.einsum("Qia,Qjb->iajb", Bov, Bov) - iajb)) < 1e-11
    ij_ab = np.einsum("Qij,Qab->ijab", Boo, Bvv)
    y = np.einsum("mnls,mi,nj,la,sb->ijab", ao, Co, Co, Cv, Cv, optimize=True)
    assert np.max(np.abs(ij_ab - y)) < 1e-11


def test_orthogonal_invariance():
    b, C, _ = random_ao(4, 8, 2)
    Boo, Bov, Bvv = df_ao_to_mo(b, C, 3)
    lhs = np.sum(Boo ** 2) + 2.0 * np.sum(Bov ** 2) + np.sum(Bvv ** 2)
    assert abs(lhs - np.sum(b ** 2)) < 1e-11 * np.sum(b ** 2)


def random_fit(nbf, naux, seed):
    """Positive definite Coulomb metric J and a 3-center tensor Pqp[mu,nu,P] symmetric in (mu, nu)."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((naux, naux))
    J = A @ A.T / naux + 0.5 * np.eye(naux)
    P = rng.standard_normal((nbf, nbf, naux))
    return J, P + np.transpose(P, (1, 0, 2))


def test_df_fit_pins():
    J, P = random_fit(6, 11, 5)
    b = df_fit(J, P)
    # sum_Q b b = Pqp J^-1 Pqp, evaluated with a linear solve instead of the square root
    four = np.einsum("Qmn,Qls->mnls", b, b)
    X = np.linalg.solve(J, P.reshape(36, 11).T)                     # J^-1 Pqp^T  [P, (l,s)]
    ref = (P.reshape(36, 11) @ X).reshape(6, 6, 6, 6)
    assert np.max(np.abs(four - ref)) < 1e-11 * np.max(np.abs(ref))
    # J = 1: b[Q,m,n] = Pqp[m,n,Q]
    assert np.max(np.abs(df_fit(np.eye(11), P) - np.transpose(P, (2, 0, 1)))) < 1e-14
    # the square root itself: b for J and J^2 relate through Jh(J^2) = Jh(J)^2
    b2 = df_fit(J @ J, P)
    d, U = np.linalg.eigh(J)
    Jm = (U / d) @ U.T                                               # J^-1
    assert np.max(np.abs(b2 - np.einsum("pqP,PQ->Qpq", P, Jm))) < 1e-11 * np.max(np.abs(b2))
