"""Writes tests/golden/*.npz: small known-answer fixtures for the DF-RCCSD(T) path.

The expected outputs come from oracle/spinorb.py, a brute-force spin-orbital CCSD/(T)
written from the textbook equations (Stanton-Gauss CCSD as tabulated by Crawford &
Schaefer; spin-orbital (T) of Raghavachari et al.).  It shares no code with the
closed-shell restatement in oracle/rccsd.py / oracle/triples.py, so these fixtures are an
independent check of both the oracle and the HIP path.

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle.spinorb import SpinOrbitalCC, closed_to_spin  # noqa: E402
from helpers import random_df, random_amplitudes  # noqa: E402


def make(name, o, v, naux, seed, scale):
    eo, ev, Boo, Bov, Bvv = random_df(o, v, naux, seed, scale)
    T1, T2 = random_amplitudes(o, v, seed)
    so = SpinOrbitalCC(Boo, Bov, Bvv, eo, ev)
    t1, t2 = closed_to_spin(T1, T2)
    E_rand = so.energy(t1, t2)
    r1, r2 = so.numerators(t1, t2)
    num1 = r1[0::2, 0::2]                 # (i alpha, a alpha)
    num2 = r2[0::2, 1::2, 0::2, 1::2]     # (i alpha, j beta, a alpha, b beta)
    Et_rand = so.triples(t1, t2)
    # MP2: one update from T = 0
    z1, z2 = np.zeros_like(t1), np.zeros_like(t2)
    m1, m2 = so.update(z1, z2)
    E_mp2 = so.energy(m1, m2)
    Ecc, c1, c2, nit = so.solve(m1, m2, tol=1e-13)
    Et_cc = so.triples(c1, c2)
    np.savez(os.path.join(HERE, name + ".npz"), eps_o=eo, eps_v=ev, Boo=Boo, Bov=Bov, Bvv=Bvv, T1=T1, T2=T2,
             E_rand=E_rand, num1=num1, num2=num2, Et_rand=Et_rand, E_mp2=E_mp2, E_ccsd=Ecc,
             T1_ccsd=c1[0::2, 0::2], T2_ccsd=c2[0::2, 1::2, 0::2, 1::2], Et_ccsd=Et_cc)
    print(f"{name}: o={o} v={v} Q={naux} E_mp2={E_mp2:.12f} E_ccsd={Ecc:.12f} (T)={Et_cc:.12e} iters={nit}")


if __name__ == "__main__":
    make("tiny_o2v3", 2, 3, 7, 11, 0.1)
    make("tiny_o3v4", 3, 4, 10, 12, 0.08)
