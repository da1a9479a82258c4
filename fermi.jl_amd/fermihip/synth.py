"""Seeded synthetic DF inputs of the shapes of the benchmark configurations.

There is no integral front end on this path (the DF tensors are the path's input, handed
over from Fermi's IntegralHelper), so workloads are generated from a seed:

* :func:`molecular_like` -- closed-shell DF tensors with the structure real-orbital DF
  tensors have: B[Q,p,q] = B[Q,q,p], a decaying auxiliary spectrum, a positive
  "orbital self-density" component on the diagonal B[Q,p,p] (Coulomb-like (pp|qq) > 0)
  and small overlap densities off the diagonal.  The off-diagonal scale is set so that
  the MP2 correlation energy is about -0.05 Eh per occupied orbital (water-like per-pair
  magnitudes), which gives CCSD iteration counts like a small molecule's (~15-25).
* :func:`water_cluster_like` -- the same structure with a geometry: n waters on a 2.8 Angstrom
  cubic lattice (positions jittered from a seed), 4 active occupied, 19 virtual orbitals and
  84 auxiliary functions per water (cc-pVDZ / cc-pVDZ-RI, frozen core).  Overlap densities
  between orbitals of different waters decay exponentially with their distance, and the fit
  coefficients decay with the distance of the auxiliary function from the density, so pair
  amplitudes fall off with the inter-water distance as in a real cluster.  Returns the orbital
  sites, from which the occupied pair distances follow.
* :func:`triples_config` -- the (T)-only synthetic tensors: eps_occ ~ U(-1.5,-0.3),
  eps_vir ~ U(0.1,3.0), B[Q,p,q] ~ N(0,1) exp(-Q/300) (symmetrised in p,q as real
  orbitals require), T2 = MP2 amplitudes from those denominators, T1 = 0 or N(0,1e-3).
* :func:`scf_like` -- an AO-basis closed-shell SCF problem for the DF-RHF path: overlap S (unit
  diagonal, positive definite), core Hamiltonian H with a spread diagonal, fitted AO tensor
  b[Q,mu,nu] symmetric in (mu, nu) and scaled so that J and K do not grow with the size.
"""
import numpy as np

__all__ = ["molecular_like", "water_cluster_like", "pair_distances", "triples_config", "mp2_amplitudes", "scf_like", "SHAPES",
           "WATERS"]

# (o, v, naux) of the benchmark configurations (cc-pVDZ / cc-pVDZ-RI water clusters, frozen core)
SHAPES = {
    "water1": (5, 19, 84),
    "water4": (16, 76, 336),
    "water10": (40, 190, 840),
    "triples64": (64, 320, 1200),
}


# number of waters of the cluster configurations
WATERS = {"water1": 1, "water4": 4, "water10": 10}
WATER_OCC, WATER_VIR, WATER_AUX = 4, 19, 84
BOHR_PER_ANGSTROM = 1.0 / 0.52917721092
# valence orbital energies of water (cc-pVDZ RHF, Eh) and a ladder of virtual energies
_EPS_OCC = np.array([-1.34, -0.70, -0.57, -0.49])
_EPS_VIR = np.array([0.18, 0.26, 0.79, 0.85, 1.16, 1.20, 1.25, 1.36, 1.45, 1.69, 1.87, 1.92, 2.45, 2.82, 3.33, 3.46,
                     3.90, 4.25, 4.60])


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


def cluster_sites(nw, seed=0, spacing_angstrom=2.8):
    """Positions (bohr) of nw waters: the nw grid points of a cubic lattice closest to its centre,
    each moved by a seeded jitter of 0.2 bohr (standing in for the random rigid rotations)."""
    rng = np.random.default_rng(seed)
    m = int(np.ceil(nw ** (1.0 / 3.0))) + 1
    g = np.array([(x, y, z) for x in range(m) for y in range(m) for z in range(m)], dtype=float)
    d2 = np.round(((g - g.mean(0)) ** 2).sum(1), 6)
    order = np.lexsort((g[:, 2], g[:, 1], g[:, 0], d2))
    pts = g[order[:nw]] * spacing_angstrom * BOHR_PER_ANGSTROM
    return pts + rng.normal(0.0, 0.2, pts.shape)


def water_cluster_like(nw, seed=0, e_mp2_per_occ=-0.05, coulomb=0.15, kappa=0.5, rho=3.0):
    """DF tensors of a synthetic (H2O)nw cluster: returns eps_o, eps_v, Boo, Bov, Bvv, sites_o, pos
    (sites_o[i] = water of occupied orbital i, pos = water positions in bohr).  Orbitals and
    auxiliary functions are grouped by water (water-major order).  B[Q,p,q] = s G[Q,p,q]
    exp(-kappa r_pq) / (1 + (r_Q,pq / rho)^2) with G a symmetrised normal tensor with a decaying
    spectrum per water, r_pq the distance of the waters of p and q and r_Q,pq the distance of the
    auxiliary function's water from the midpoint of the density; the diagonal carries a positive
    self-density as in molecular_like.  s is calibrated to e_mp2_per_occ."""
    rng = np.random.default_rng(seed)
    pos = cluster_sites(nw, seed)
    o, v, naux = WATER_OCC * nw, WATER_VIR * nw, WATER_AUX * nw
    n = o + v
    eps_o = np.concatenate([_EPS_OCC + rng.normal(0.0, 0.02, WATER_OCC) for _ in range(nw)])
    eps_v = np.concatenate([_EPS_VIR * rng.uniform(0.95, 1.05) + rng.normal(0.0, 0.03, WATER_VIR)
                            for _ in range(nw)])
    site = np.concatenate([np.repeat(np.arange(nw), WATER_OCC), np.repeat(np.arange(nw), WATER_VIR)])
    qsite = np.repeat(np.arange(nw), WATER_AUX)
    qloc = np.tile(np.arange(WATER_AUX), nw)
    decay = np.exp(-qloc / (WATER_AUX / 3.0))
    r_pq = np.linalg.norm(pos[site][:, None, :] - pos[site][None, :, :], axis=-1)        # n x n
    mid = 0.5 * (pos[site][:, None, :] + pos[site][None, :, :])                           # n x n x 3
    amp = np.empty((naux, n, n))
    for w in range(nw):
        r_q = np.linalg.norm(mid - pos[w], axis=-1)
        amp[qsite == w] = (np.exp(-kappa * r_pq) / (1.0 + (r_q / rho) ** 2))[None]
    amp *= decay[:, None, None]
    G = _sym(rng.standard_normal((naux, n, n))) * amp
    dens = np.abs(rng.standard_normal((naux, n))) * amp[:, np.arange(n), np.arange(n)]
    Gov = G[:, :o, o:]
    T = mp2_amplitudes(eps_o, eps_v, Gov)
    ovov = np.tensordot(Gov, Gov, axes=(0, 0))
    e1 = float(np.sum((2 * T - np.transpose(T, (0, 1, 3, 2))) * np.transpose(ovov, (0, 2, 1, 3))))
    s = (e_mp2_per_occ * o / e1) ** 0.25
    B = s * G
    idx = np.arange(n)
    B[:, idx, idx] += coulomb * dens
    Boo = np.ascontiguousarray(B[:, :o, :o])
    Bov = np.ascontiguousarray(B[:, :o, o:])
    Bvv = np.ascontiguousarray(B[:, o:, o:])
    return eps_o, eps_v, Boo, Bov, Bvv, site[:o].copy(), pos


def pair_distances(sites_o, pos):
    """Occupied pair distances d[i,j] (bohr) between the waters of orbitals i and j."""
    p = pos[sites_o]
    return np.linalg.norm(p[:, None, :] - p[None, :, :], axis=-1)


def scf_like(nbf, naux, seed=0):
    """(S, H, b) of a synthetic closed-shell SCF problem."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((nbf, nbf))
    S = np.eye(nbf) + 0.05 * (A + A.T)
    lo = np.linalg.eigvalsh(S).min()
    if lo < 0.3:
        S = S + (abs(lo) + 0.3) * np.eye(nbf)
    d = np.sqrt(np.diag(S))
    S = S / d[:, None] / d[None, :]
    B = rng.standard_normal((nbf, nbf))
    H = np.diag(np.linspace(-5.0, 2.0, nbf)) + 0.1 * (B + B.T)
    G = rng.standard_normal((naux, nbf, nbf))
    b = 0.06 * np.sqrt(420.0 / (nbf * naux)) * (G + np.transpose(G, (0, 2, 1)))
    return S, H, b
