"""Shared seeded inputs for the tests (tiny random DF problems)."""
import numpy as np


def random_df(o, v, naux, seed, scale=0.3):
    """Symmetric random DF tensors B[Q,p,q] = B[Q,q,p] and ordered orbital energies."""
    rng = np.random.default_rng(seed)
    n = o + v
    G = rng.standard_normal((naux, n, n))
    G = scale * (G + np.transpose(G, (0, 2, 1)))
    eps_o = np.sort(rng.uniform(-1.5, -0.3, o))
    eps_v = np.sort(rng.uniform(0.1, 3.0, v))
    return eps_o, eps_v, G[:, :o, :o].copy(), G[:, :o, o:].copy(), G[:, o:, o:].copy()


def random_amplitudes(o, v, seed, s1=0.1, s2=0.1):
    """Unconverged amplitudes with the closed-shell symmetry T2[i,j,a,b] = T2[j,i,b,a]."""
    rng = np.random.default_rng(seed + 1000)
    T1 = s1 * rng.standard_normal((o, v))
    T2 = s2 * rng.standard_normal((o, o, v, v))
    T2 = T2 + np.transpose(T2, (1, 0, 3, 2))
    return T1, T2


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
