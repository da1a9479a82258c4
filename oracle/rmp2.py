"""CPU oracle (test infrastructure only): FP64 restatement of Fermi.jl's DF-MP2 energy
``RMP2_energy(ints::IntegralHelper{T,<:AbstractDFERI,RHFOrbitals}, ::RMP2Algorithm)``.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import
this module, and only as the checker.  The product path (``libfermihip.so``) never calls it.

Pins (``tests/test_oracle_mp2.py``):
  * the golden fixtures' ``E_mp2``, which ``tests/golden/make_golden.py`` takes from the
    brute-force spin-orbital code (``oracle/spinorb.py``: one Jacobi update from T = 0);
  * the textbook spin-orbital MP2 sum 1/4 sum |<ij||ab>|^2 / D_ijab (Szabo & Ostlund eq. 6.72)
    over the antisymmetrised integrals of ``oracle/spinorb.py``;
  * the one-pair closed form o = v = 1: E = (ia|ia)^2 / (2 (e_i - e_a)).
"""
import numpy as np

__all__ = ["rmp2_energy_df"]


def rmp2_energy_df(Bov, eps_o, eps_v):
    """RMP2a.jl:91-144.  Bov[Q,i,a] = moints["BOV"], eps_o = moints["Fii"], eps_v = moints["Faa"].

    Bvo = permutedims(BOV, (1,3,2)) (:92); for each pair i <= j (:114-120) the v x v block
    Bab = Bvo[:,:,i]^T Bvo[:,:,j] = (ia|jb) (:122), summed as Bab[a,b] (2 Bab[a,b] - Bab[b,a]) /
    (e_i + e_j - e_a - e_b) (:124-131), with weight 2 for i != j (:132-133)."""
    Bvo = np.transpose(Bov, (0, 2, 1))
    o, v = len(eps_o), len(eps_v)
    Emp2 = 0.0
    for i in range(o):
        Bi = Bvo[:, :, i]
        for j in range(i, o):
            Bab = Bi.T @ Bvo[:, :, j]
            D = eps_o[i] + eps_o[j] - eps_v[:, None] - eps_v[None, :]
            E = float(np.sum(Bab * (2.0 * Bab - Bab.T) / D))
            Emp2 += (2.0 if i != j else 1.0) * E
    return Emp2
