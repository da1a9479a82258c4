"""CPU oracle (test infrastructure only): FP64 restatement of Fermi.jl's DF fit of
``compute_ERI!`` (AtomicIntegrals.jl:31-41), the AO -> MO DF transforms ``compute_BOO!`` /
``compute_BOV!`` / ``compute_BVV!`` (DFERI.jl:15-69) and the RHF orbital energies of
``compute_F!`` (Fock.jl:81-98).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import this
module, and only as the checker.  The product path (``libfermihip.so``) never calls it.

Arrays are indexed like the Julia arrays: ``bAO[Q,mu,nu]``, ``C[mu,p]``, ``B[Q,p,q]``.

Pins (``tests/test_oracle_dferi.py``):
  * DF fit: the fitted four-index tensor sum_Q b b equals Pqp J^-1 Pqp computed with a linear
    solve (no square root); J = 1 gives b = Pqp permuted; the fit with J^2 in place of J applies
    J^-1 (the square root squared);
  * C = identity: every block is the matching slice of bAO;
  * four-index consistency: sum_Q Bov[Q,i,a] Bov[Q,j,b] equals the AO four-index tensor
    (mu nu|la si) = sum_Q b b transformed one index at a time, on an AO tensor that is not symmetric
    in (mu, nu), so a swapped occupied/virtual coefficient fails;
  * orthogonal invariance: for orthogonal C, no dropped orbitals and symmetric b, the Frobenius
    norm |Boo|^2 + 2 |Bov|^2 + |Bvv|^2 equals |b|^2.
"""
import numpy as np

__all__ = ["active_spaces", "df_ao_to_mo", "rhf_eps", "df_fit"]


def df_fit(J, Pqp):
    """b[Q,p,q] = Pqp[p,q,P] Jh[P,Q] with Jh = real(J^(-1/2)) (AtomicIntegrals.jl:35-40), the
    symmetric inverse square root of the positive definite Coulomb metric."""
    d, U = np.linalg.eigh(0.5 * (J + J.T))
    Jh = (U / np.sqrt(d)) @ U.T
    return np.einsum("pqP,PQ->Qpq", Pqp, Jh, optimize=True)


def active_spaces(nbf, ndocc, drop_occ=0, drop_vir=0):
    """o = (1+core):ndocc, v = (ndocc+1):(nbf-inac) (DFERI.jl:20,39-40), as 0-based ranges."""
    return range(drop_occ, ndocc), range(ndocc, nbf - drop_vir)


def rhf_eps(eps, ndocc, drop_occ=0, drop_vir=0):
    """Fii = eps[o], Faa = eps[v] (Fock.jl:93-94)."""
    o, v = active_spaces(len(eps), ndocc, drop_occ, drop_vir)
    return np.asarray(eps)[list(o)], np.asarray(eps)[list(v)]


def df_ao_to_mo(bAO, C, ndocc, drop_occ=0, drop_vir=0):
    """(BOO, BOV, BVV): Bij = b[Q,mu,nu] Co[mu,i] Co[nu,j] (DFERI.jl:25-27),
    Bia = b Co[mu,i] Cv[nu,a] (:46-48), Bab = b Cv[mu,a] Cv[nu,b] (:64-66)."""
    nbf = C.shape[0]
    o, v = active_spaces(nbf, ndocc, drop_occ, drop_vir)
    Co, Cv = C[:, list(o)], C[:, list(v)]
    Bij = np.einsum("Qmn,mi,nj->Qij", bAO, Co, Co, optimize=True)
    Bia = np.einsum("Qmn,mi,na->Qia", bAO, Co, Cv, optimize=True)
    Bab = np.einsum("Qmn,ma,nb->Qab", bAO, Cv, Cv, optimize=True)
    return Bij, Bia, Bab
