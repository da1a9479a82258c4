"""Pins of the DF-MP2 oracle (oracle/rmp2.py, a restatement of RMP2a.jl:91-144).

Each check is independent of the restatement's formula: the golden fixtures (from the
brute-force spin-orbital CCSD code), the textbook spin-orbital MP2 sum over antisymmetrised
integrals, and the one-pair closed form.  A dropped pair weight, a sign error or a transposed
exchange integral fails one of them.
"""
import os

import numpy as np
import pytest

from oracle.rmp2 import rmp2_energy_df
from oracle.spinorb import SpinOrbitalCC
from helpers import random_df


def test_golden_fixtures():
    here = os.path.join(os.path.dirname(__file__), "golden")
    for name in ("tiny_o2v3", "tiny_o3v4"):
        g = np.load(os.path.join(here, name + ".npz"))
        assert abs(rmp2_energy_df(g["Bov"], g["eps_o"], g["eps_v"]) - g["E_mp2"]) < 1e-13


@pytest.mark.parametrize("o,v,Q,seed", [(1, 2, 4, 0), (2, 3, 6, 1), (3, 4, 9, 2), (3, 2, 5, 3)])
def test_spin_orbital_textbook_sum(o, v, Q, seed):
    """E = 1/4 sum_{IJAB} |<IJ||AB>|^2 / D_IJAB over spin orbitals (Szabo & Ostlund eq. 6.72)."""
    eo, ev, Boo, Bov, Bvv = random_df(o, v, Q, seed)
    so = SpinOrbitalCC(Boo, Bov, Bvv, eo, ev)
    no = so.no
    g = so.g[:no, :no, no:, no:]
    E_so = 0.25 * float(np.sum(g * g / so.Dijab))
    assert abs(rmp2_energy_df(Bov, eo, ev) - E_so) < 1e-12 * max(1.0, abs(E_so))


def test_one_pair_closed_form():
    rng = np.random.default_rng(4)
    Bov = rng.standard_normal((7, 1, 1))
    eo, ev = np.array([-0.7]), np.array([0.4])
    x = float(np.sum(Bov[:, 0, 0] ** 2))                  # (ia|ia)
    assert abs(rmp2_energy_df(Bov, eo, ev) - x * x / (2.0 * (eo[0] - ev[0]))) < 1e-14


def test_orbital_relabelling_invariance():
    """The energy is a sum over all pairs: relabelling occupied or virtual orbitals (with their
    energies) leaves it unchanged."""
    eo, ev, Boo, Bov, Bvv = random_df(4, 5, 8, 11)
    po, pv = np.array([2, 0, 3, 1]), np.array([4, 1, 0, 3, 2])
    E = rmp2_energy_df(Bov, eo, ev)
    assert abs(rmp2_energy_df(Bov[:, po][:, :, pv], eo[po], ev[pv]) - E) < 1e-13
