"""Analytic two-electron integrals on the GPU (dg_eri_*, McMurchie-Davidson)
against the oracle's restatement of the reference's THO integrals
(or_eri_J, src/integrals.cpp:389-468): two independent algorithms, FP64.

- J = sum_kl P_kl (ij|kl) on H2O/STO-3G, CH4/6-31G and CO with a Cartesian d
  shell (the (dd|dd) class), to 1e-12 of the largest element;
- one normalised s primitive: (ss|ss) = 2 sqrt(a / pi_t) F_0(1e-8) in closed
  form (the reference's truncated pi and Boys clamp);
- the packed layout is the reference's teindex, and the 8-fold symmetry holds;
- the SCF with hartree_evaluation = two_electron_integrals against the
  oracle's SCF on the same path, to 1e-8 Ha.
"""
import math

import numpy as np
import pytest

from dftcxx_amd import gridlib as G
from dftcxx_amd import molecule as M

from systems import random_density, rel_err, with_d_shell

pytestmark = pytest.mark.gpu


def _co_d():
    return with_d_shell(M.Molecule.from_angstrom(["C", "O"], [[0, 0, 0], [0, 0, 1.13]], "sto3g"))


@pytest.mark.parametrize("make", [lambda: M.h2o("sto3g"), lambda: M.ch4("p631"), _co_d], ids=["h2o", "ch4", "co_d"])
def test_coulomb_vs_oracle(make):
    from oracle import oracle as O
    mol = make()
    print(f"ERI J on {mol.name}: {mol.nbf()} functions", flush=True)
    e = G.Eri(mol)
    n = mol.nbf()
    assert e.nints == (n * (n + 1) // 2) * (n * (n + 1) // 2 + 1) // 2
    P = random_density(n, 13)
    J = e.coulomb(P)
    Jo = O.eri_J(mol, P)
    assert rel_err(J, Jo) < 1e-12
    assert np.array_equal(J, J.T)


def test_single_primitive_closed_form():
    a = 0.8
    mol = M.Molecule([1], [[0.0, 0.0, 0.0]], {1: [(0, [a], [1.0])]}, "s")
    ints = G.Eri(mol).integrals()
    pi_t = 3.14159265359
    f0 = 1.0 - 1e-8 / 3.0 + 1e-16 / 10.0  # F_0 at the clamped argument 1e-8
    assert abs(ints[0] - 2.0 * math.sqrt(a / pi_t) * f0) < 1e-14


def test_teindex_layout_and_symmetry():
    mol = M.h2o("sto3g")
    e = G.Eri(mol)
    ints = e.integrals()
    n = mol.nbf()
    rng = np.random.default_rng(3)
    P = np.zeros((n, n))
    for _ in range(40):
        i, j, k, l = rng.integers(0, n, 4)
        v = ints[G.teindex(i, j, k, l)]
        for p in [(j, i, k, l), (i, j, l, k), (k, l, i, j), (l, k, j, i)]:
            assert ints[G.teindex(*p)] == v
        # a single (k, l) entry of P picks out (ij|kl) in J
        P[:] = 0.0
        P[k, l] = 1.0
        J = e.coulomb(P)
        assert abs(J[i, j] - v) < 1e-14 * max(1.0, abs(v))


@pytest.mark.parametrize("step", [G.STEP_HOST, G.STEP_DEVICE], ids=["host_step", "device_step"])
def test_scf_two_electron_integrals_vs_oracle(step):
    from oracle import oracle as O
    mol = M.h2o("sto3g")
    og = O.Grid(mol, 30, 110, 6)
    en_o, terms_o, P_o = og.scf(threshold=1e-6, eri=True)
    en_g, res, P_g = G.scf_run(mol, 30, 110, 6, threshold=1e-6, hartree=G.HARTREE_ERI, density_step=step)
    assert res["converged"] == 1 and len(en_g) == len(en_o)
    assert np.abs(en_g - en_o).max() < 1e-8
    assert abs(res["e_j"] - terms_o["J"]) < 1e-8 and abs(res["e_xc"] - terms_o["xc"]) < 1e-8
    assert np.abs(P_g - P_o).max() < 1e-7
