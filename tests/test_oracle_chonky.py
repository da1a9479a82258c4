"""Pins of the conventional transform oracle (oracle/chonky.py, a restatement of Chonky.jl:1-131):
identity coefficients, agreement with the DF route for an AO tensor of DF form, and the 8-fold
permutational symmetry."""
import numpy as np

from oracle.chonky import ao2mo_4index, chonky_block
from oracle.dferi import df_ao_to_mo


def df_form(naux, nbf, seed):
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((naux, nbf, nbf))
    b = b + np.transpose(b, (0, 2, 1))
    C = rng.standard_normal((nbf, nbf))
    return b, np.einsum("Qmn,Qrs->mnrs", b, b), C


def test_identity_coefficients():
    _, eri, _ = df_form(4, 6, 0)
    I = np.eye(6)
    assert np.array_equal(ao2mo_4index(eri, I[:, :2], I[:, 2:5], I[:, 1:3], I[:, 4:]), eri[:2, 2:5, 1:3, 4:])


def test_blocks_equal_the_df_route():
    b, eri, C = df_form(7, 8, 1)
    nd, core, inac = 4, 1, 1
    Boo, Bov, Bvv = df_ao_to_mo(b, C, nd, core, inac)
    want = {"OOOO": np.einsum("Qij,Qkl->ijkl", Boo, Boo), "OOOV": np.einsum("Qij,Qka->ijka", Boo, Bov),
            "OOVV": np.einsum("Qij,Qab->ijab", Boo, Bvv), "OVOV": np.einsum("Qia,Qjb->iajb", Bov, Bov),
            "OVVV": np.einsum("Qia,Qbc->iabc", Bov, Bvv), "VVVV": np.einsum("Qab,Qcd->abcd", Bvv, Bvv)}
    for key, w in want.items():
        got = chonky_block(eri, C, nd, key, core, inac)
        assert np.max(np.abs(got - w)) < 1e-10 * np.max(np.abs(w)), key


def test_eightfold_symmetry():
    _, eri, C = df_form(5, 6, 2)
    mo = ao2mo_4index(eri, C, C, C, C)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]:
        assert np.max(np.abs(mo - np.transpose(mo, perm))) < 1e-10 * np.max(np.abs(mo))
