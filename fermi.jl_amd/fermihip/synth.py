"""Seeded synthetic DF inputs of the shapes of the benchmark configurations.

There is no integral front end on this path (the DF tensors are the path's input, handed
over from Fermi's IntegralHelper), so workloads are generated from a seed:

* :func:`molecular_like` -- closed-shell DF tensors with the structure real-orbital DF
  tensors have: B[Q,p,q] = B[Q,q,p], a decaying auxiliary spectrum, a positive
  "orbital self-density" component on the diagonal B[Q,p,p] (Coulomb-like (pp|qq) > 0)
  and small overlap densities off the diagonal.  The off-diagonal scale is set so that
  the MP2 correlation energy is about -0.05 Eh per occupied orbital (water-like per-pair
  magnitudes), which gives CCSD iteration counts like a small molecule's (~15-25).
* :func:`triples_config` -- the (T)-only synthetic tensors: eps_occ ~ U(-1.5,-0.3),
  eps_vir ~ U(0.1,3.0), B[Q,p,q] ~ N(0,1) exp(-Q/300) (symmetrised in p,q as real
  orbitals require), T2 = MP2 amplitudes from those denominators, T1 = 0 or N(0,1e-3).
"""
import numpy as np

__all__ = ["molecular_like", "triples_config", "mp2_amplitudes", "SHAPES"]

# (o, v, naux) of the benchmark configurations (cc-pVDZ / cc-pVDZ-RI water clusters, frozen core)
SHAPES = {
    "water1": (5, 19, 84),
    "water4": (16, 76, 336),
    "water10": (40, 190, 840),
    "triples64": (64, 320, 1200),
}


def _sym(G):
    return (G + np.transpose(G, (0, 2, 1))) / np.sqrt(2.0)


def mp2_amplitudes(eps_o, eps_v, Bov):
    ovov = np.tensordot(Bov, Bov, axes=(0, 0))          # (ia|jb)
    D = eps_o[:, None, None, None] + eps_o[None, :, None, None] - eps_v[None, None, :, None] - eps_v[None, None, None, :]
    return np.transpose(ovov, (0, 2, 1, 3)) / D


def molecular_like(o, v, naux, seed=0, e_mp2_per_occ=-0.05, coulomb=0.15):
    rng = np.random.default_rng(seed)
    n = o + v
    eps_o = np.sort(rng.uniform(-1.5, -0.3, o))
    eps_v = np.sort(rng.uniform(0.1, 3.0, v))
    decay = np.exp(-np.arange(naux) / (naux / 3.0))[:, None, None]
    G = _sym(rng.standard_normal((naux, n, n))) * decay
    dens = np.abs(rng.standard_normal((naux, n))) * decay[:, :, 0]
    # MP2 energy is quartic in the off-diagonal scale s: calibrate s from s = 1
    Gov = G[:, :o, o:]
    T = mp2_amplitudes(eps_o, eps_v, Gov)
    ovov = np.tensordot(Gov, Gov, axes=(0, 0))
    e1 = float(np.sum((2 * T - np.transpose(T, (0, 1, 3, 2))) * np.transpose(ovov, (0, 2, 1, 3))))
    s = (e_mp2_per_occ * o / e1) ** 0.25
    B = s * G
    idx = np.arange(n)
    B[:, idx, idx] += coulomb * dens / np.sqrt(max(naux, 84) / 84.0)
    Boo = np.ascontiguousarray(B[:, :o, :o])
    Bov = np.ascontiguousarray(B[:, :o, o:])
    Bvv = np.ascontiguousarray(B[:, o:, o:])
    return eps_o, eps_v, Boo, Bov, Bvv


def triples_config(o=64, v=320, naux=1200, seed=0, t1_scale=0.0):
    rng = np.random.default_rng(seed)
    n = o + v
    eps_o = rng.uniform(-1.5, -0.3, o)
    eps_v = rng.uniform(0.1, 3.0, v)
    B = _sym(rng.standard_normal((naux, n, n))) * np.exp(-np.arange(naux) / 300.0)[:, None, None]
    Boo = np.ascontiguousarray(B[:, :o, :o])
    Bov = np.ascontiguousarray(B[:, :o, o:])
    Bvv = np.ascontiguousarray(B[:, o:, o:])
    T2 = mp2_amplitudes(eps_o, eps_v, Bov)
    T1 = t1_scale * rng.standard_normal((o, v)) if t1_scale > 0 else np.zeros((o, v))
    return eps_o, eps_v, Boo, Bov, Bvv, T1, T2
