"""Pins of the (T) oracle (oracle/triples.py, restatements of ijk.jl, ijk2.jl and abc.jl)."""
import os

import numpy as np
import pytest

from oracle import rccsd as R
from oracle import triples as TT
from oracle.spinorb import SpinOrbitalCC, closed_to_spin
from helpers import random_df, random_amplitudes, relerr


@pytest.mark.parametrize("o,v,Q,seed", [(2, 3, 6, 0), (3, 4, 9, 1), (3, 5, 7, 2)])
def test_ijk_matches_spin_orbital_triples(o, v, Q, seed):
    eo, ev, Boo, Bov, Bvv = random_df(o, v, Q, seed)
    m = R.build_moints(Boo, Bov, Bvv)
    T1, T2 = random_amplitudes(o, v, seed)
    so = SpinOrbitalCC(Boo, Bov, Bvv, eo, ev)
    t1, t2 = closed_to_spin(T1, T2)
    Eso = so.triples(t1, t2)
    assert abs(TT.rccsdpt_ijk(T1, T2, m, eo, ev) - Eso) < 1e-12 * abs(Eso)


@pytest.mark.parametrize("o,v,Q,seed", [(3, 4, 9, 3), (4, 5, 8, 4)])
def test_three_reference_algorithms_agree(o, v, Q, seed):
    eo, ev, Boo, Bov, Bvv = random_df(o, v, Q, seed)
    m = R.build_moints(Boo, Bov, Bvv)
    T1, T2 = random_amplitudes(o, v, seed)
    e1 = TT.rccsdpt_ijk(T1, T2, m, eo, ev)
    assert abs(TT.rccsdpt_ijk2(T1, T2, m, eo, ev) - e1) < 1e-12 * abs(e1)
    assert abs(TT.rccsdpt_abc(T1, T2, m, eo, ev, faithful=False) - e1) < 1e-12 * abs(e1)


def test_abc_reference_operands_differ():
    """abc.jl:128,131 contract T2_c where the permutation pattern needs T2_a; restated
    literally the abc algorithm does not reproduce ijk.jl (recorded reference defect)."""
    o, v, Q = 3, 4, 9
    eo, ev, Boo, Bov, Bvv = random_df(o, v, Q, 5)
    m = R.build_moints(Boo, Bov, Bvv)
    T1, T2 = random_amplitudes(o, v, 5)
    e1 = TT.rccsdpt_ijk(T1, T2, m, eo, ev)
    assert abs(TT.rccsdpt_abc(T1, T2, m, eo, ev, faithful=True) - e1) > 1e-6 * abs(e1)


def test_scaling_laws_and_relabelling():
    o, v, Q = 3, 5, 8
    eo, ev, Boo, Bov, Bvv = random_df(o, v, Q, 6)
    m = R.build_moints(Boo, Bov, Bvv)
    _, T2 = random_amplitudes(o, v, 6)
    T1 = np.zeros((o, v))
    e1 = TT.rccsdpt_ijk(T1, T2, m, eo, ev)
    # T1 = 0: E_T(lambda T2) = lambda^2 E_T(T2)
    assert abs(TT.rccsdpt_ijk(T1, 1.7 * T2, m, eo, ev) - 1.7 ** 2 * e1) < 1e-12 * abs(e1) * 3
    # integrals scaled by mu (B by sqrt(mu)) at fixed amplitudes: E_T -> mu^2 E_T
    mu = 0.6
    ms = R.build_moints(np.sqrt(mu) * Boo, np.sqrt(mu) * Bov, np.sqrt(mu) * Bvv)
    T1r, T2r = random_amplitudes(o, v, 16)
    a = TT.rccsdpt_ijk(T1r, T2r, m, eo, ev)
    assert abs(TT.rccsdpt_ijk(T1r, T2r, ms, eo, ev) - mu ** 2 * a) < 1e-12 * abs(a) * 3
    # relabelling within o and within v
    po, pv = np.array([1, 2, 0]), np.array([3, 0, 4, 2, 1])
    m2 = R.build_moints(Boo[:, po][:, :, po], Bov[:, po][:, :, pv], Bvv[:, pv][:, :, pv])
    b = TT.rccsdpt_ijk(T1r[po][:, pv], T2r[po][:, po][:, :, pv][:, :, :, pv], m2, eo[po], ev[pv])
    assert abs(b - a) < 1e-12 * abs(a)


def test_diagonal_triples_vanish():
    """i=j=k has weight 2-1-1=0 (ijk.jl:133); a=b=c terms vanish identically (X=6WV, Y=Z=3V)."""
    o, v, Q = 3, 4, 7
    eo, ev, Boo, Bov, Bvv = random_df(o, v, Q, 8)
    m = R.build_moints(Boo, Bov, Bvv)
    T1, T2 = random_amplitudes(o, v, 8)
    full = TT.rccsdpt_ijk(T1, T2, m, eo, ev)
    nodiag = TT.rccsdpt_ijk(T1, T2, m, eo, ev, triples=[t for t in TT.triple_list(o) if not (t[0] == t[1] == t[2])])
    assert full == pytest.approx(nodiag, rel=1e-14, abs=0)


def test_golden_triples():
    here = os.path.join(os.path.dirname(__file__), "golden")
    for name in ("tiny_o2v3", "tiny_o3v4"):
        g = np.load(os.path.join(here, name + ".npz"))
        m = R.build_moints(g["Boo"], g["Bov"], g["Bvv"])
        assert abs(TT.rccsdpt_ijk(g["T1"], g["T2"], m, g["eps_o"], g["eps_v"]) - g["Et_rand"]) < 1e-12 * abs(g["Et_rand"])
        e = TT.rccsdpt_ijk(g["T1_ccsd"], g["T2_ccsd"], m, g["eps_o"], g["eps_v"])
        assert abs(e - g["Et_ccsd"]) < 1e-12
