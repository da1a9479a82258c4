"""Pins of the CCSD oracle (oracle/rccsd.py, a restatement of RCCSDHelper.jl / RCCSDa.jl / DIIS.jl).

Every check is independent of the restatement's own formulas: a brute-force spin-orbital
CCSD from the textbook equations, the DF-MP2 closed form, the conventional-ladder identity,
closed-shell symmetry, orbital relabelling and the Pulay closed form of DIIS.
"""
import numpy as np
import pytest

from oracle import rccsd as R
from oracle.spinorb import SpinOrbitalCC, closed_to_spin
from helpers import random_df, random_amplitudes, relerr


@pytest.mark.parametrize("o,v,Q,seed", [(1, 2, 5, 0), (2, 3, 7, 1), (3, 4, 9, 2), (2, 5, 6, 3)])
def test_numerators_match_spin_orbital_ccsd(o, v, Q, seed):
    """Closed-shell T1/T2 numerators = (ia)/(ia jb) blocks of the spin-orbital CCSD equations at
    random unconverged amplitudes (a dropped term, wrong sign or transposed index fails this)."""
    eo, ev, Boo, Bov, Bvv = random_df(o, v, Q, seed)
    m = R.build_moints(Boo, Bov, Bvv)
    T1, T2 = random_amplitudes(o, v, seed)
    so = SpinOrbitalCC(Boo, Bov, Bvv, eo, ev)
    t1, t2 = closed_to_spin(T1, T2)
    r1, r2 = so.numerators(t1, t2)
    assert relerr(R.cc_update_T1(T1, T2, m), r1[0::2, 0::2]) < 1e-12
    assert relerr(R.cc_update_T2(T1, T2, m), r2[0::2, 1::2, 0::2, 1::2]) < 1e-12
    assert abs(R.cc_update_energy(T1, T2, m) - so.energy(t1, t2)) < 1e-12 * max(1.0, abs(so.energy(t1, t2)))


def test_golden_fixtures():
    import os
    here = os.path.join(os.path.dirname(__file__), "golden")
    for name in ("tiny_o2v3", "tiny_o3v4"):
        g = np.load(os.path.join(here, name + ".npz"))
        m = R.build_moints(g["Boo"], g["Bov"], g["Bvv"])
        assert relerr(R.cc_update_T1(g["T1"], g["T2"], m), g["num1"]) < 1e-12
        assert relerr(R.cc_update_T2(g["T1"], g["T2"], m), g["num2"]) < 1e-12
        assert abs(R.cc_update_energy(g["T1"], g["T2"], m) - g["E_rand"]) < 1e-12
        T1, T2 = R.mp2_guess(m, g["eps_o"], g["eps_v"])
        assert abs(R.cc_update_energy(T1, T2, m) - g["E_mp2"]) < 1e-13
        res = R.rccsd_solve(m, g["eps_o"], g["eps_v"], T1, T2, R.Options(diis_prec="double"))
        assert res["converged"]
        assert abs(res["correlation"] - g["E_ccsd"]) < 1e-10
        assert relerr(res["T2"], g["T2_ccsd"]) < 1e-7


def test_mp2_closed_form():
    """From T=0 one update gives T1=0 and T2=(ia|jb)/D2; the energy is the DF-MP2 formula
    sum (ia|jb)(2(ia|jb)-(ib|ja))/D (RMP2a.jl:112-138)."""
    o, v, Q = 4, 7, 15
    eo, ev, Boo, Bov, Bvv = random_df(o, v, Q, 5, 0.1)
    m = R.build_moints(Boo, Bov, Bvv)
    z1, z2 = np.zeros((o, v)), np.zeros((o, o, v, v))
    n1, n2 = R.update_amp(z1, z2, m, eo, ev)
    assert np.max(np.abs(n1)) == 0.0
    T1, T2 = R.mp2_guess(m, eo, ev)
    assert relerr(n2, T2) < 1e-14
    iajb = np.einsum("Qia,Qjb->iajb", Bov, Bov)
    D = eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]
    E = np.sum(iajb * (2 * iajb - np.transpose(iajb, (0, 3, 2, 1))) / D)
    assert abs(R.cc_update_energy(T1, T2, m) - E) < 1e-13


def test_df_ladder_equals_conventional_ladder():
    o, v, Q = 3, 6, 11
    eo, ev, Boo, Bov, Bvv = random_df(o, v, Q, 7)
    m = R.build_moints(Boo, Bov, Bvv)
    T1, T2 = random_amplitudes(o, v, 7)
    assert relerr(R.df_ladder(T1, T2, Bvv), R.conv_ladder(T1, T2, m)) < 1e-13
    assert relerr(R.cc_update_T2(T1, T2, m, df=True), R.cc_update_T2(T1, T2, m, df=False)) < 1e-13


def test_closed_shell_symmetry_and_relabelling():
    o, v, Q = 3, 5, 10
    eo, ev, Boo, Bov, Bvv = random_df(o, v, Q, 9)
    m = R.build_moints(Boo, Bov, Bvv)
    T1, T2 = random_amplitudes(o, v, 9)
    n2 = R.cc_update_T2(T1, T2, m)
    assert relerr(n2, np.transpose(n2, (1, 0, 3, 2))) < 1e-13   # RCCSDHelper.jl:190 keeps T2[i,j,a,b]=T2[j,i,b,a]
    # relabel orbitals within o and within v: the update commutes with the relabelling
    po, pv = np.array([2, 0, 1]), np.array([4, 1, 3, 0, 2])
    m2 = R.build_moints(Boo[:, po][:, :, po], Bov[:, po][:, :, pv], Bvv[:, pv][:, :, pv])
    T1p, T2p = T1[po][:, pv], T2[po][:, po][:, :, pv][:, :, :, pv]
    a1, a2 = R.update_amp(T1, T2, m, eo, ev)
    b1, b2 = R.update_amp(T1p, T2p, m2, eo[po], ev[pv])
    assert relerr(b1, a1[po][:, pv]) < 1e-13
    assert relerr(b2, a2[po][:, po][:, :, pv][:, :, :, pv]) < 1e-13


def test_solve_reaches_spin_orbital_fixed_point():
    o, v, Q = 3, 6, 12
    eo, ev, Boo, Bov, Bvv = random_df(o, v, Q, 4, 0.08)
    m = R.build_moints(Boo, Bov, Bvv)
    T1, T2 = R.mp2_guess(m, eo, ev)
    res = R.rccsd_solve(m, eo, ev, T1, T2, R.Options(diis_prec="single"))
    assert res["converged"]
    so = SpinOrbitalCC(Boo, Bov, Bvv, eo, ev)
    t1, t2 = closed_to_spin(T1, T2)
    Eso, _, _, _ = so.solve(t1, t2, tol=1e-13)
    assert abs(res["correlation"] - Eso) < 1e-9
    # DIIS precision changes the trajectory, not the fixed point
    res2 = R.rccsd_solve(m, eo, ev, T1, T2, R.Options(diis_prec="double"))
    assert abs(res2["correlation"] - res["correlation"]) < 1e-9


def test_nonconvergence_is_not_an_error_and_eps_clamp():
    o, v, Q = 2, 4, 8
    eo, ev, Boo, Bov, Bvv = random_df(o, v, Q, 2, 0.1)
    m = R.build_moints(Boo, Bov, Bvv)
    T1, T2 = R.mp2_guess(m, eo, ev)
    res = R.rccsd_solve(m, eo, ev, T1, T2, R.Options(cc_max_iter=2))
    assert not res["converged"] and res["niter"] == 2          # RCCSDa.jl:179-182, no throw
    # thresholds below eps(Float64) are clamped to eps unless precision_override (RCCSDa.jl:139-146)
    res = R.rccsd_solve(m, eo, ev, T1, T2, R.Options(cc_e_conv=1e-20, cc_max_rms=1e-20, cc_max_iter=200))
    assert res["converged"]


def test_diis_closed_form_and_errors():
    dm = R.DIISManager(3, np.float64, np.float32)
    x = np.random.default_rng(0).random(10)
    dm.push(x, x)
    assert len(dm) == 1
    with pytest.raises(IndexError):           # BoundsError in test/test_diis.jl:1-9
        dm.extrapolate()
    # two vectors: c1 = (e2.e2 - e1.e2)/|e1-e2|^2, c2 = 1-c1 (Pulay)
    dm = R.DIISManager(3, np.float64, np.float64)
    rng = np.random.default_rng(1)
    v1, v2, e1, e2 = rng.random(6), rng.random(6), rng.random(6), rng.random(6)
    dm.push(v1, e1)
    dm.push(v2, e2)
    c1 = (e2 @ e2 - e1 @ e2) / ((e1 - e2) @ (e1 - e2))
    c2 = 1 - c1
    assert np.allclose(dm.extrapolate(), c1 * v1 + c2 * v2, atol=1e-13)
    assert np.allclose(dm.extrapolate(add_res=True), c1 * (v1 + e1) + c2 * (v2 + e2), atol=1e-13)
    # when full, the vector with the largest error norm is dropped (DIIS.jl:42-47)
    dm = R.DIISManager(2, np.float64, np.float64)
    dm.push(np.ones(3), 5 * np.ones(3))
    dm.push(2 * np.ones(3), np.ones(3))
    dm.push(3 * np.ones(3), 2 * np.ones(3))
    assert [float(v[0]) for v in dm.vecs] == [2.0, 3.0]
