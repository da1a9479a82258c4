"""GPU parity of the DF fit (fh_df_fit, AtomicIntegrals.jl:31-41) and the AO -> MO DF transform
(fh_set_df_ao / fh_get_df, through the C ABI) against the oracle's restatement of DFERI.jl:15-69 (pinned on the CPU by the identity-coefficient case,
four-index consistency and orthogonal invariance).  FP64 MFMA GEMMs: 1e-12 relative.  Downstream,
a context made from the AO tensor gives the same CCSD update and DF-MP2 energy as one made from
the oracle's MO tensors.
"""
import sys

import numpy as np
import pytest

from oracle.dferi import df_ao_to_mo, df_fit, rhf_eps
from oracle.rmp2 import rmp2_energy_df
from helpers import relerr

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_ao(naux, nbf, seed, symmetric=True):
    rng = np.random.default_rng(seed)
    b = 0.2 * rng.standard_normal((naux, nbf, nbf))
    if symmetric:
        b = b + np.transpose(b, (0, 2, 1))
    C, _ = np.linalg.qr(rng.standard_normal((nbf, nbf)))
    eps = np.sort(rng.uniform(-2.0, 3.0, nbf))
    return b, C, eps


@pytest.mark.parametrize("naux,nbf,nd,core,inac,sym", [(7, 9, 3, 0, 0, False), (30, 20, 6, 1, 2, False),
                                                        (84, 60, 12, 2, 0, True), (336, 96, 20, 4, 0, True)])
def test_ao_to_mo_matches_oracle(naux, nbf, nd, core, inac, sym):
    import fermihip
    say(f"ao_to_mo naux={naux} nbf={nbf} ndocc={nd} drop_occ={core} drop_vir={inac}")
    b, C, eps = random_ao(naux, nbf, nbf, sym)
    c = fermihip.Context(0)
    c.set_df_ao(eps, C, b, nd, core, inac)
    ref = df_ao_to_mo(b, C, nd, core, inac)
    for got, want in zip(c.get_df(), ref):
        assert got.shape == want.shape
        assert relerr(got, want) < 1e-12
    eo, ev = rhf_eps(eps, nd, core, inac)
    e = c.mp2_energy()
    e_ref = rmp2_energy_df(ref[1], eo, ev)
    assert abs(e - e_ref) < 1e-12 * max(1.0, abs(e_ref))


def test_ao_context_gives_the_mo_context_update():
    """CCSD update and energies from fh_set_df_ao equal those from fh_set_df with the oracle's blocks."""
    import fermihip
    from helpers import random_amplitudes
    say("ao_to_mo downstream update")
    naux, nbf, nd, core, inac = 40, 22, 5, 1, 3
    b, C, eps = random_ao(naux, nbf, 3)
    eo, ev = rhf_eps(eps, nd, core, inac)
    Boo, Bov, Bvv = df_ao_to_mo(b, C, nd, core, inac)
    ca = fermihip.Context(0)
    ca.set_df_ao(eps, C, b, nd, core, inac)
    cm = fermihip.Context(0)
    cm.set_df(eo, ev, Boo, Bov, Bvv)
    T1, T2 = random_amplitudes(nd - core, nbf - nd - inac, 3, 0.02, 0.02)
    a1, a2 = ca.update_amp(T1, T2)
    m1, m2 = cm.update_amp(T1, T2)
    assert relerr(a1, m1) < 1e-11 and relerr(a2, m2) < 1e-11
    assert abs(ca.energy(T1, T2) - cm.energy(T1, T2)) < 1e-12 * max(1.0, abs(cm.energy(T1, T2)))


def test_ao_to_mo_errors():
    import fermihip
    say("ao_to_mo errors")
    b, C, eps = random_ao(5, 8, 4)
    c = fermihip.Context(0)
    with pytest.raises(fermihip.FermiException, match="frozen orbitals"):
        c.set_df_ao(eps, C, b, 3, 3, 0)
    with pytest.raises(fermihip.FermiException, match="inactive virtual"):
        c.set_df_ao(eps, C, b, 3, 0, 5)
    with pytest.raises(fermihip.FermiException, match="occupied orbital energy"):
        c.set_df_ao(eps[::-1].copy(), C, b, 3, 0, 0)


@pytest.mark.parametrize("nbf,naux,seed", [(3, 5, 0), (10, 25, 1), (30, 84, 2)])
def test_df_fit_matches_oracle(nbf, naux, seed):
    import fermihip
    from test_oracle_dferi import random_fit
    say(f"df_fit nbf={nbf} naux={naux}")
    J, P = random_fit(nbf, naux, seed)
    b = fermihip.Context(0).df_fit(J, P)
    assert relerr(b, df_fit(J, P)) < 1e-11


def test_df_fit_errors():
    import fermihip
    say("df_fit errors")
    J = -np.eye(4)
    with pytest.raises(fermihip.FermiException, match="positive definite"):
        fermihip.Context(0).df_fit(J, np.ones((2, 2, 4)))


@pytest.mark.parametrize("nbf,nd,core,inac,seed", [(5, 2, 0, 0, 0), (14, 5, 1, 2, 1), (24, 5, 0, 0, 2)])
def test_ao2mo_4index_matches_oracle(nbf, nd, core, inac, seed):
    """Conventional transform (Chonky.jl:1-131): the full MO tensor and every block."""
    import fermihip
    from oracle.chonky import ao2mo_4index, chonky_block
    say(f"ao2mo_4index nbf={nbf}")
    rng = np.random.default_rng(seed)
    eri = rng.standard_normal((nbf, nbf, nbf, nbf))
    C = rng.standard_normal((nbf, nbf))
    c = fermihip.Context(0)
    assert relerr(c.ao2mo_4index(eri, C, C, C, C), ao2mo_4index(eri, C, C, C, C)) < 1e-12
    o, v = slice(core, nd), slice(nd, nbf - inac)
    blk = {"O": C[:, o], "V": C[:, v]}
    for key in ("OOOO", "OOOV", "OOVV", "OVOV", "OVVV", "VVVV"):
        got = c.ao2mo_4index(eri, *[blk[ch] for ch in key])
        assert relerr(got, chonky_block(eri, C, nd, key, core, inac)) < 1e-12, key
