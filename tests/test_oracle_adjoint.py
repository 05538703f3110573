"""Pins of the E_xc adjoint restated in oracle/adjoint.py (CPU only).

The derivatives are checked against central finite differences of the oracle's
own E_xc pipeline (set_density -> correct_densities -> xc) and against the
closed identities dE/dP|_c = 2 c XC and sum_A dE/dR_A = 0 (the grid and the
Becke partition move with the atoms, so E_xc is translation invariant).
"""
import numpy as np
import pytest

from dftcxx_amd import molecule as M
from oracle import adjoint as AD
from oracle import oracle as O
from systems import core_guess, random_density, with_d_shell

NRAD, NANG, LMAX = 20, 26, 4


def _case(name):
    mol = M.h2o() if name == "h2o" else with_d_shell(M.ch4("sto3g"))
    P = core_guess(mol, *O.one_electron(mol))
    if name != "h2o":
        P = P + 0.05 * random_density(mol.nbf(), seed=5)
    return mol, P


def test_basis_values_match_oracle_amplitudes():
    mol = with_d_shell(M.co("sto3g"))
    g = O.Grid(mol, NRAD, NANG, LMAX)
    xyz, _, _ = g.points()
    phi, dphi, centre = AD.basis_values(mol, xyz)
    ref = g.amplitudes()
    assert np.abs(phi - ref).max() < 1e-14 * np.abs(ref).max()
    # gradient of each function against a central difference of the values
    h = 1e-5
    for d in range(3):
        dx = np.zeros(3)
        dx[d] = h
        fp = AD.basis_values(mol, xyz + dx)[0]
        fm = AD.basis_values(mol, xyz - dx)[0]
        fd = (fp - fm) / (2 * h)
        assert np.abs(fd - dphi[:, :, d]).max() < 1e-6 * max(1.0, np.abs(dphi).max())


def test_becke_weight_gradient_by_finite_differences():
    mol = M.h2o()
    g = O.Grid(mol, 8, 14, 2)
    xyz, w, wb = g.points()
    owner = np.arange(len(w)) // (8 * 14)
    wb2, dwb = AD.becke_weight_gradient(mol, xyz, owner)
    assert np.abs(wb2 - wb).max() < 1e-14
    h = 1e-5
    for a in range(mol.natoms):
        for d in range(3):
            vals = []
            for sgn in (1, -1):
                m2 = M.Molecule(mol.Z, mol.xyz.copy(), mol.basis, mol.name)
                m2.xyz[a, d] += sgn * h
                x2 = xyz.copy()
                x2[owner == a, d] += sgn * h   # the points of cell a move with their atom
                vals.append(AD.becke_weight_gradient(m2, x2, owner)[0])
            fd = (vals[0] - vals[1]) / (2 * h)
            assert np.abs(fd - dwb[:, a, d]).max() < 1e-6


@pytest.mark.parametrize("name", ["h2o", "ch4_d"])
def test_dE_dP_by_finite_differences(name):
    mol, P = _case(name)
    exc, dEdP, _ = AD.xc_adjoint(mol, NRAD, NANG, LMAX, P)
    n = mol.nbf()
    rng = np.random.default_rng(3)
    h = 1e-5
    for _ in range(4):
        i, j = rng.integers(0, n, 2)
        E = np.zeros((n, n))
        E[i, j] += 0.5
        E[j, i] += 0.5
        ep = AD.exc_at(mol, NRAD, NANG, LMAX, P + h * E)[0]
        em = AD.exc_at(mol, NRAD, NANG, LMAX, P - h * E)[0]
        fd = (ep - em) / (2 * h)
        an = 0.5 * (dEdP[i, j] + dEdP[j, i])
        assert abs(fd - an) < 1e-7 * max(1.0, abs(an)), (i, j, fd, an)


def test_dE_dP_identity_with_c_held():
    mol, P = _case("h2o")
    g = O.Grid(mol, NRAD, NANG, LMAX)
    g.set_density(P)
    S = g.calculate_density()
    g.correct_densities()
    XC, _ = g.xc()
    c = float(np.sum(mol.Z)) / S
    _, dEdP, _ = AD.xc_adjoint(mol, NRAD, NANG, LMAX, P, hold_c=True)
    assert np.abs(dEdP - 2 * c * XC).max() < 1e-12 * np.abs(XC).max()


@pytest.mark.parametrize("name", ["h2o", "ch4_d"])
def test_forces_by_finite_differences(name):
    mol, P = _case(name)
    _, _, grad = AD.xc_adjoint(mol, NRAD, NANG, LMAX, P)
    _, _, grad_c = AD.xc_adjoint(mol, NRAD, NANG, LMAX, P, hold_c=True)
    # translation invariance
    assert np.abs(grad.sum(0)).max() < 1e-10 * max(1.0, np.abs(grad).max())
    h = 1e-4
    g = O.Grid(mol, NRAD, NANG, LMAX)
    g.set_density(P)
    S0 = g.calculate_density()
    g.correct_densities()
    g.xc()
    rho = g.densities()
    _, wv = g.xc_parts()
    Ztot = float(np.sum(mol.Z))
    k = float(wv @ rho) / Ztot
    c = Ztot / S0
    for a in range(mol.natoms):
        for d in range(3):
            out = []
            for sgn in (1, -1):
                m2 = M.Molecule(mol.Z, mol.xyz.copy(), mol.basis, mol.name)
                m2.xyz[a, d] += sgn * h
                out.append(AD.exc_at(m2, NRAD, NANG, LMAX, P))
            fd = (out[0][0] - out[1][0]) / (2 * h)
            assert abs(fd - grad[a, d]) < 5e-8 * max(1.0, abs(grad[a, d])), (a, d, fd, grad[a, d])
            # c held: the difference is k c dS/dR
            dS = (out[0][1] - out[1][1]) / (2 * h)
            assert abs(grad_c[a, d] - (grad[a, d] + k * c * dS)) < 5e-8 * max(1.0, abs(grad_c[a, d]))
