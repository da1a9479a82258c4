"""CPU oracle (test infrastructure only): FP64 restatement of Fermi.jl's conventional ("Chonky")
4-index AO -> MO transforms, ``compute_ERI!`` and ``compute_OOOO!`` ... ``compute_VVVV!``
(Chonky.jl:1-131).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import this
module, and only as the checker.  The product path (``libfermihip.so``) never calls it.

Arrays are indexed like the Julia arrays: ``eri[mu,nu,rho,sigma]`` (chemists' order), ``C[mu,p]``.

Pins (``tests/test_oracle_chonky.py``):
  * identity coefficients give slices of the AO tensor;
  * for an AO tensor of density-fitted form eri = sum_Q b b, every block equals the same block
    assembled from the DF transform of ``oracle/dferi.py`` (sum_Q B B), an independent route;
  * the 8-fold permutational symmetry of real-orbital integrals survives the transform.
"""
import numpy as np

from .dferi import active_spaces

__all__ = ["ao2mo_4index", "chonky_block"]


def ao2mo_4index(eri, C1, C2, C3, C4):
    """MOERI[i,j,k,l] = AOERI[mu,nu,rho,sigma] C1[mu,i] C2[nu,j] C3[rho,k] C4[sigma,l] (Chonky.jl:7)."""
    return np.einsum("mnrs,mi,nj,rk,sl->ijkl", eri, C1, C2, C3, C4, optimize=True)


def chonky_block(eri, C, ndocc, key, drop_occ=0, drop_vir=0):
    """compute_<key>! for key in OOOO, OOOV, OOVV, OVOV, OVVV, VVVV (Chonky.jl:11-131): the letters
    pick the occupied (drop_occ+1 : ndocc) or virtual (ndocc+1 : nbf-drop_vir) column block of C."""
    o, v = active_spaces(C.shape[0], ndocc, drop_occ, drop_vir)
    blk = {"O": C[:, list(o)], "V": C[:, list(v)]}
    return ao2mo_4index(eri, *[blk[ch] for ch in key])
