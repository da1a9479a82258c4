"""Pins of the DF-RHF oracle (oracle/rhf.py, a restatement of RHFa.jl:67-223 and RHFHelper.jl):
the conventional Fock build on the four-index tensor, a two-function system whose orbital is
fixed by symmetry (closed-form energy), and the stationarity conditions of the converged SCF."""
import numpy as np
import pytest
from scipy.linalg import expm

from oracle import rhf as RH
from helpers import random_scf


def test_df_fock_equals_conventional_fock():
    S, H, b = random_scf(9, 13, 0)
    rng = np.random.default_rng(1)
    Co = rng.standard_normal((9, 3))
    D = Co @ Co.T
    eri = np.einsum("Qmn,Qrs->mnrs", b, b)
    assert np.max(np.abs(RH.build_fock_df(H, D, b) - RH.build_fock_conv(H, D, eri))) < 1e-12


def test_symmetric_two_function_closed_form():
    """H2-like: two equivalent functions, one doubly occupied orbital g = (1,1)/sqrt(2(1+s)),
    E = 2 h_gg + (gg|gg) + Vnuc with h_gg = (H11 + H12)/(1+s)."""
    s, h11, h12, Vnuc = 0.4, -1.1, -0.9, 0.7
    S = np.array([[1.0, s], [s, 1.0]])
    H = np.array([[h11, h12], [h12, h11]])
    rng = np.random.default_rng(2)
    naux = 5
    b = np.zeros((naux, 2, 2))
    x, y = rng.uniform(0.2, 0.5, naux), rng.uniform(0.0, 0.2, naux)
    b[:, 0, 0] = b[:, 1, 1] = x                                  # mirror symmetry 1 <-> 2
    b[:, 0, 1] = b[:, 1, 0] = y
    g = np.array([1.0, 1.0]) / np.sqrt(2.0 * (1.0 + s))
    hgg = (h11 + h12) / (1.0 + s)
    Bg = np.einsum("Qmn,m,n->Q", b, g, g)
    E_closed = 2.0 * hgg + float(Bg @ Bg) + Vnuc
    # (the core guess is already the exact orbital here: the ODA step of RHFa.jl:171-175 then
    # divides 0 by 0, as the reference does)
    for diis in (True, False):
        r = RH.rhf(S, H, b, 1, Vnuc, RH.Options(scf_guess="gwh", diis=diis))
        assert r["converged"]
        assert abs(r["energy"] - E_closed) < 1e-12


@pytest.mark.parametrize("seed,guess", [(1, "gwh"), (2, "core"), (3, "gwh")])
def test_converged_stationarity(seed, guess):
    nbf, naux, nd = 14, 30, 4
    S, H, b = random_scf(nbf, naux, seed)
    r = RH.rhf(S, H, b, nd, 0.0, RH.Options(scf_guess=guess))
    assert r["converged"]
    C, eps = r["C"], r["eps"]
    D = C[:, :nd] @ C[:, :nd].T
    F = RH.build_fock_df(H, D, b)
    assert np.max(np.abs(C.T @ S @ C - np.eye(nbf))) < 1e-10
    assert np.max(np.abs(F @ D @ S - S @ D @ F)) < 1e-5
    assert np.max(np.abs(np.diag(C.T @ F @ C) - eps)) < 1e-5
    assert abs(RH.rhf_energy(D, H, F) - r["energy"]) < 1e-9
    # energy rises under small occupied-virtual rotations
    rng = np.random.default_rng(seed)
    for _ in range(3):
        K = np.zeros((nbf, nbf))
        X = 1e-3 * rng.standard_normal((nd, nbf - nd))
        K[:nd, nd:] = X
        K[nd:, :nd] = -X.T
        Cr = C @ expm(K)                                             # orthogonal rotation
        Dr = Cr[:, :nd] @ Cr[:, :nd].T
        assert RH.rhf_energy(Dr, H, RH.build_fock_df(H, Dr, b)) > r["energy"]
