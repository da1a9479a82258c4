"""GPU parity of the DF-RHF SCF (fh_rhf, through the C ABI) against the oracle's restatement of
RHFa.jl:67-223 / RHFHelper.jl (pinned on the CPU by the conventional Fock build, a closed-form
two-function system and the stationarity of converged solutions).  Same guess, same ODA / DIIS
decisions: the iteration count must match and the energy agree to 1e-10 relative; orbitals are
compared through the density (eigenvector signs are arbitrary)."""
import sys

import numpy as np
import pytest

from oracle import rhf as RH
from oracle.dferi import df_ao_to_mo, rhf_eps
from oracle.rmp2 import rmp2_energy_df
from helpers import random_scf

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def density(C, nd):
    return C[:, :nd] @ C[:, :nd].T


@pytest.mark.parametrize("nbf,naux,nd,seed,guess", [(14, 30, 4, 1, "gwh"), (14, 30, 4, 2, "core"), (14, 30, 4, 3, "gwh"),
                                                     (40, 90, 9, 2, "gwh"), (96, 336, 20, 1, "core")])
def test_rhf_matches_oracle(nbf, naux, nd, seed, guess):
    import fermihip
    say(f"rhf nbf={nbf} naux={naux} ndocc={nd} seed={seed} guess={guess}")
    S, H, b = random_scf(nbf, naux, seed)
    ref = RH.rhf(S, H, b, nd, 1.25, RH.Options(scf_guess=guess))
    r = fermihip.RHF(S, H, b, nd, 1.25, alg=fermihip.RHFhip(), options=fermihip.ScfOptions(scf_guess=guess))
    assert r.converged == ref["converged"] and r.niter == ref["niter"]
    assert abs(r.energy - ref["energy"]) < 1e-10 * abs(ref["energy"])
    assert np.max(np.abs(r.eps - ref["eps"])) < 1e-8
    assert np.max(np.abs(density(r.C, nd) - density(ref["C"], nd))) < 1e-8
    assert np.max(np.abs(r.C.T @ S @ r.C - np.eye(nbf))) < 1e-10


def test_rhf_two_function_closed_form():
    import fermihip
    say("rhf two-function closed form")
    s, h11, h12, Vnuc = 0.4, -1.1, -0.9, 0.7
    S = np.array([[1.0, s], [s, 1.0]])
    H = np.array([[h11, h12], [h12, h11]])
    rng = np.random.default_rng(2)
    b = np.zeros((5, 2, 2))
    x, y = rng.uniform(0.2, 0.5, 5), rng.uniform(0.0, 0.2, 5)
    b[:, 0, 0] = b[:, 1, 1] = x
    b[:, 0, 1] = b[:, 1, 0] = y
    g = np.array([1.0, 1.0]) / np.sqrt(2.0 * (1.0 + s))
    Bg = np.einsum("Qmn,m,n->Q", b, g, g)
    E_closed = 2.0 * (h11 + h12) / (1.0 + s) + float(Bg @ Bg) + Vnuc
    # the guess is already the exact orbital, so every DIIS error vector is exactly zero and the
    # DIIS matrix singular (the reference's B \ resid throws there too): run with ODA alone
    r = fermihip.RHF(S, H, b, 1, Vnuc, options=fermihip.ScfOptions(diis=False))
    assert r.converged and abs(r.energy - E_closed) < 1e-12


def test_rhf_nonconvergence_and_errors():
    import fermihip
    say("rhf non-convergence and errors")
    S, H, b = random_scf(14, 30, 1)
    ref = RH.rhf(S, H, b, 4, 0.0, RH.Options(scf_max_iter=3))
    r = fermihip.RHF(S, H, b, 4, 0.0, options=fermihip.ScfOptions(scf_max_iter=3))
    assert not r.converged and r.niter == 3                 # not an error (RHFa.jl:213-218)
    assert abs(r.energy - ref["energy"]) < 1e-10 * abs(ref["energy"])
    bad = S.copy()
    bad[0, 1] = bad[1, 0] = 1.5                              # not positive definite
    with pytest.raises(fermihip.FermiException, match="positive definite"):
        fermihip.RHF(bad, H, b, 4)
    with pytest.raises(fermihip.FermiException):
        fermihip.RHF(S, H, b, 4, options=fermihip.ScfOptions(ndiis=12))


def test_rhf_to_mp2_chain():
    """RHF -> AO->MO transform (one frozen core) -> DF-MP2, on the device and in the oracle."""
    import fermihip
    say("rhf -> ao_to_mo -> mp2 chain")
    nbf, naux, nd, core = 24, 50, 6, 1
    S, H, b = random_scf(nbf, naux, 1)
    ref = RH.rhf(S, H, b, nd, 0.0)
    assert ref["converged"]
    r = fermihip.RHF(S, H, b, nd, 0.0)
    c = fermihip.Context(0)
    c.set_df_ao(r.eps, r.C, b, nd, core, 0)
    e_gpu = c.mp2_energy()
    eo, ev = rhf_eps(ref["eps"], nd, core, 0)
    e_ref = rmp2_energy_df(df_ao_to_mo(b, ref["C"], nd, core, 0)[1], eo, ev)
    assert abs(e_gpu - e_ref) < 1e-9 * abs(e_ref)
