"""GPU parity of the E_xc adjoint (dE_xc/dP and dE_xc/dR) through the C ABI.

The device adjoint is compared with the numpy restatement in oracle/adjoint.py
(itself pinned by finite differences on the CPU), and checked directly against
central finite differences of the library's own E_xc, the identity
dE/dP|_c = 2 c XC and translation invariance (sum_A dE/dR_A = 0).
"""
import numpy as np
import pytest

from dftcxx_amd import gridlib as G
from dftcxx_amd import molecule as M

from systems import core_guess, random_density, with_d_shell

pytestmark = pytest.mark.gpu


def _density(mol, extra=0.0):
    from oracle import oracle as O
    P = core_guess(mol, *O.one_electron(mol))
    if extra:
        P = P + extra * random_density(mol.nbf(), seed=5)
    return P


def _lib_adjoint(mol, nrad, nang, lmax, P, hold_c, precision=G.PREC_FP64):
    g = G.MolecularGrid(mol, nrad, nang, lmax, precision=precision)
    g.set_density(P)
    g.correct_densities()
    XC, exc = g.calculate_exchange_correlation()
    dP, grad = g.xc_adjoint(hold_c=hold_c)
    return g, XC, exc, dP, grad


@pytest.mark.parametrize("name", ["h2o", "ch4_d", "benzene_sto3g", "c10h22"])
@pytest.mark.parametrize("hold_c", [False, True])
def test_adjoint_matches_oracle(name, hold_c):
    from oracle import adjoint as AD
    if name == "h2o":
        mol, grid = M.h2o(), (20, 26, 4)
        P = _density(mol)
    elif name == "ch4_d":
        mol, grid = with_d_shell(M.ch4("sto3g")), (20, 26, 4)
        P = _density(mol, 0.05)
    elif name == "c10h22":
        # 134 functions in 9 tiles: the streamed T kernel with its per-block tile
        # mask, the listed dE/dP contraction and the shell reach cut all take part
        mol, grid = M.alkane(10), (10, 26, 2)
        P = _density(mol)
    else:
        mol, grid = M.benzene("sto3g"), (16, 26, 4)
        P = _density(mol)
    print(f"[case] adjoint {name} hold_c={hold_c}", flush=True)
    exc_o, dP_o, grad_o = AD.xc_adjoint(mol, *grid, P, hold_c=hold_c)
    _, _, exc, dP, grad = _lib_adjoint(mol, *grid, P, hold_c)
    assert abs(exc - exc_o) < 1e-11 * abs(exc_o)
    assert np.abs(dP - dP_o).max() < 1e-11 * np.abs(dP_o).max()
    assert np.abs(grad - grad_o).max() < 1e-10 * max(1.0, np.abs(grad_o).max())


def test_adjoint_identities_benzene_full_grid():
    mol = M.benzene("p631")
    # a density matrix invariant under every symmetry operation of the molecule
    # (equal diagonal values for equivalent functions; a core-Hamiltonian guess
    # can occupy one function of a degenerate pair and break the symmetry)
    P = 0.2 * np.eye(mol.nbf())
    print("[case] adjoint benzene 6-31G 75x302", flush=True)
    g, XC, exc, dP_c, grad_c = _lib_adjoint(mol, 75, 302, 8, P, True)
    dP, grad = g.xc_adjoint(hold_c=False)
    # c held: dE/dP = 2 c XC (c from the electron count before the correction)
    g2 = G.MolecularGrid(mol, 75, 302, 8)
    g2.set_density(P)
    S = g2.correct_densities()
    c = float(np.sum(mol.Z)) / S
    assert np.abs(dP_c - 2 * c * XC).max() < 1e-12 * np.abs(XC).max()
    # c variable: dE/dP gains -2 c (Lambda / Z) S_grid, Lambda = sum w v rho
    Sg = g.get_amplitudes().T @ (g.get_weights()[:, None] * g.get_amplitudes())
    _, wv = g.get_xc_parts()
    lam = float(wv @ g.get_densities())
    expect = 2 * c * lam / float(np.sum(mol.Z))
    assert np.abs(dP_c - dP - expect * Sg).max() < 1e-10 * np.abs(dP).max()
    for gr in (grad, grad_c):
        assert np.abs(gr.sum(0)).max() < 1e-9 * np.abs(gr).max()
    # the symmetry operations shared by D6h and the octahedral Lebedev grid map
    # the gradient onto itself: inversion (atom k -> k+3) and the mirror y -> -y
    # (k -> -k); the 60-degree rotations are not grid symmetries
    tol = 1e-10 * np.abs(grad).max()
    for base in (0, 6):
        for k in range(6):
            g, gi = grad[base + k], grad[base + (k + 3) % 6]
            gm = grad[base + (6 - k) % 6]
            assert np.abs(gi + g).max() < tol
            assert np.abs(gm - g * [1, -1, 1]).max() < tol
    assert np.abs(grad[:, 2]).max() < tol


def test_forces_by_finite_differences_on_device():
    mol = M.h2o()
    P = _density(mol)
    nrad, nang, lmax = 30, 50, 6
    print("[case] adjoint FD h2o 30x50", flush=True)
    _, _, _, _, grad = _lib_adjoint(mol, nrad, nang, lmax, P, False)
    h = 1e-4
    for a, d in ((0, 1), (1, 0), (1, 1), (2, 2)):
        e = []
        for sgn in (1, -1):
            m2 = M.Molecule(mol.Z, mol.xyz.copy(), mol.basis, mol.name)
            m2.xyz[a, d] += sgn * h
            g = G.MolecularGrid(m2, nrad, nang, lmax)
            g.set_density(P)
            g.correct_densities()
            e.append(g.calculate_exchange_correlation()[1])
        fd = (e[0] - e[1]) / (2 * h)
        assert abs(fd - grad[a, d]) < 5e-8 * max(1.0, abs(grad[a, d])), (a, d, fd, grad[a, d])


def test_adjoint_in_reduced_precision_mode_is_fp64():
    mol = M.h2o("p631")
    P = _density(mol)
    print("[case] adjoint mixed-mode handle", flush=True)
    _, _, _, dP6, g6 = _lib_adjoint(mol, 30, 50, 6, P, False)
    _, _, _, dPm, gm = _lib_adjoint(mol, 30, 50, 6, P, False, precision=G.PREC_MIXED)
    assert np.abs(dPm - dP6).max() < 1e-12 * np.abs(dP6).max()
    assert np.abs(gm - g6).max() < 1e-12 * max(1.0, np.abs(g6).max())


def test_adjoint_needs_xc_first():
    mol = M.h2o()
    g = G.MolecularGrid(mol, 20, 26, 4)
    with pytest.raises(G.DgError):
        g.xc_adjoint()
    g.set_density(_density(mol))
    with pytest.raises(G.DgError):
        g.xc_adjoint()  # density not yet corrected / no XC


def test_c10h22_forces_and_dEdP_full_grid():
    # decane, 6-31G, 32 atoms x 75 x 302: forces against central differences
    # (h = 1e-4 bohr) within 1e-5 Ha/bohr, and dE/dP|_c = 2 c V_xc
    mol = M.alkane(10)
    P = _density(mol)
    nrad, nang, lmax = 75, 302, 8
    print("[case] adjoint C10H22 6-31G 75x302", flush=True)
    g, XC, exc, dP_c, _ = _lib_adjoint(mol, nrad, nang, lmax, P, True)
    dP, grad = g.xc_adjoint(hold_c=False)
    assert np.abs(grad.sum(0)).max() < 1e-9 * np.abs(grad).max()
    g2 = G.MolecularGrid(mol, nrad, nang, lmax)
    g2.set_density(P)
    c = float(np.sum(mol.Z)) / g2.correct_densities()
    del g2
    assert np.abs(dP_c - 2 * c * XC).max() < 1e-12 * np.abs(XC).max()
    h = 1e-4
    # a terminal carbon, a middle carbon and a hydrogen, one component each
    for a, d in ((0, 0), (4, 1), (12, 2)):
        e = []
        for sgn in (1, -1):
            m2 = M.Molecule(mol.Z, mol.xyz.copy(), mol.basis, mol.name)
            m2.xyz[a, d] += sgn * h
            gd = G.MolecularGrid(m2, nrad, nang, lmax)
            gd.set_density(P)
            gd.correct_densities()
            e.append(gd.calculate_exchange_correlation()[1])
            del gd
        fd = (e[0] - e[1]) / (2 * h)
        print(f"  atom {a} dir {d}: analytic {grad[a, d]:.10f} fd {fd:.10f}", flush=True)
        assert abs(fd - grad[a, d]) < 1e-5, (a, d, fd, grad[a, d])
