"""GPU parity of the DF-MP2 energy (fh_mp2_energy, through the C ABI) against the oracle's
restatement of RMP2a.jl:91-144, which the CPU tests pin against the golden fixtures, the textbook
spin-orbital MP2 sum and the one-pair closed form.  FP64 throughout: 1e-12 relative.
"""
import os
import sys

import numpy as np
import pytest

from oracle.rmp2 import rmp2_energy_df
from helpers import random_df

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mp2_ctx(eo, ev, Bov, Boo=None, Bvv=None):
    import fermihip
    c = fermihip.Context(0)
    c.set_df(eo, ev, Boo, Bov, Bvv)
    return c


@pytest.mark.parametrize("o,v,Q,seed", [(1, 1, 3, 0), (2, 3, 7, 1), (5, 19, 84, 3), (13, 17, 40, 5), (8, 30, 50, 4)])
def test_mp2_matches_oracle(o, v, Q, seed):
    say(f"mp2 o={o} v={v} Q={Q} seed={seed}")
    eo, ev, Boo, Bov, Bvv = random_df(o, v, Q, seed)
    ref = rmp2_energy_df(Bov, eo, ev)
    assert abs(mp2_ctx(eo, ev, Bov).mp2_energy() - ref) < 1e-12 * max(1.0, abs(ref))
    # a full context gives the same number, and so does the CC energy of the MP2 guess
    c = mp2_ctx(eo, ev, Bov, Boo, Bvv)
    assert abs(c.mp2_energy() - ref) < 1e-12 * max(1.0, abs(ref))
    c.guess(0, copy_out=False)
    assert abs(c.energy() - ref) < 1e-12 * max(1.0, abs(ref))


@pytest.mark.parametrize("name", ["tiny_o2v3", "tiny_o3v4"])
def test_mp2_golden(name):
    say(f"mp2 golden {name}")
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    assert abs(mp2_ctx(g["eps_o"], g["eps_v"], g["Bov"]).mp2_energy() - g["E_mp2"]) < 1e-12


def test_mp2_blocks_of_j(monkeypatch):
    """X = (ia|jb) built one block of j at a time (ragged last block) gives the one-block energy."""
    o, v, Q = 13, 60, 50
    say(f"mp2 blocks of j o={o} v={v} Q={Q}")
    eo, ev, Boo, Bov, Bvv = random_df(o, v, Q, 6)
    ref = rmp2_energy_df(Bov, eo, ev)
    c = mp2_ctx(eo, ev, Bov)
    e1 = c.mp2_energy()
    monkeypatch.setenv("FERMIHIP_MP2_BLOCK_MB", "1")   # 2 j per block: 7 blocks, the last with one j
    e7 = c.mp2_energy()
    assert abs(e1 - ref) < 1e-12 * abs(ref)
    assert abs(e7 - e1) < 1e-13 * abs(ref)


def test_mp2_water4_and_mp2_only_context():
    """o=16, v=76, Q=336; an MP2-only context (Bov alone) refuses the CC entry points."""
    import fermihip
    from fermihip import synth
    o, v, Q = synth.SHAPES["water4"]
    say(f"mp2 water4 ({o},{v},{Q})")
    eo, ev, Boo, Bov, Bvv = synth.molecular_like(o, v, Q, 1)
    ref = rmp2_energy_df(Bov, eo, ev)
    c = mp2_ctx(eo, ev, Bov)
    e = c.mp2_energy()
    assert abs(e - ref) < 1e-12 * abs(ref)
    assert c.mp2_energy() == e                           # deterministic
    with pytest.raises(fermihip.FermiException, match="Boo and Bvv"):
        c.guess(0)
    with pytest.raises(fermihip.FermiException, match="Boo and Bvv"):
        c.prepare()
    with pytest.raises(fermihip.FermiException):
        c.set_df(eo, ev, Boo, Bov, None)                 # one of the two blocks only
    m = fermihip.MOIntegrals(None, Bov, None, eo, ev, sd_energy=-76.0)
    r = fermihip.RMP2(m, alg=fermihip.RMP2hip())
    assert abs(r.correlation - ref) < 1e-12 * abs(ref) and abs(r.energy - (ref - 76.0)) < 1e-9
