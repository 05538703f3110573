"""GPU checks of the largest workload: linear C30H62, 6-31G (394 CGFs), 92 atoms.

On a reduced grid (6 x 14, lmax 2) every stage's output is compared with the
oracle: the 394-function paths (the streamed density kernel over 25 tiles,
the 2-D split contraction, FP32 contraction in mixed mode) are the ones the
full grid runs.  On the full 75 x 302, lmax 8 grid (2,083,800 points) the
checks are size-independent: exact symmetry of J and XC, the electron count,
the identity 2 tr(PJ) = sum w V rho / (2c), screened vs dense, mixed vs FP64,
dE/dP|_c = 2c XC, translation invariance of the forces and central finite
differences of E_xc (h = 1e-4 bohr, within 1e-5 Ha/bohr).
"""
import numpy as np
import pytest

from dftcxx_amd import gridlib as G
from dftcxx_amd import molecule as M

from systems import core_guess, rel_err

pytestmark = pytest.mark.gpu

_CACHE = {}


def _mol_and_guess():
    if "mol" not in _CACHE:
        mol = M.alkane(30)
        _CACHE["mol"] = mol
        _CACHE["P"] = core_guess(mol, *G.one_electron(mol))
    return _CACHE["mol"], _CACHE["P"]


def _host_build(g, P):
    g.set_density(P)
    nb = g.correct_densities()
    J = g.calculate_hartree_potential()
    XC, exc = g.calculate_exchange_correlation()
    return J, XC, exc, nb


def _device_build(g, P):
    import torch
    n = g.nbf
    dev = torch.device("cuda:0")
    dP = torch.tensor(P.T.copy(), dtype=torch.float64, device=dev)
    dJ = torch.zeros((n, n), dtype=torch.float64, device=dev)
    dX = torch.zeros((n, n), dtype=torch.float64, device=dev)
    dS = torch.zeros(4, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    g.build_device(dP.data_ptr(), dJ.data_ptr(), dX.data_ptr(), dS.data_ptr())
    torch.cuda.synchronize()
    return dJ.cpu().numpy().T.copy(), dX.cpu().numpy().T.copy(), dS.cpu().numpy().copy()


def test_c30h62_reduced_grid_against_oracle():
    from oracle import oracle as O
    mol, P = _mol_and_guess()
    assert mol.nbf() == 394 and mol.natoms == 92
    grid = (6, 14, 2)
    print("[case] C30H62 6-31G 6x14 lmax 2 vs oracle", flush=True)
    o = O.Grid(mol, *grid)
    o.set_density(P)
    o.correct_densities()
    rho_o = o.densities()
    Jo = o.hartree()
    XCo, exco = o.xc()
    for prec in (G.PREC_FP64, G.PREC_MIXED):
        g = G.MolecularGrid(mol, *grid, precision=prec)
        J, XC, exc, _ = _host_build(g, P)
        tol = 1e-10 if prec == G.PREC_FP64 else 1e-6
        print(f"  precision {prec}: J {rel_err(J, Jo):.2e} XC {rel_err(XC, XCo):.2e} "
              f"E_xc {exc - exco:.2e}", flush=True)
        assert rel_err(g.get_densities(), rho_o) < 1e-12
        assert rel_err(J, Jo) < tol
        assert rel_err(XC, XCo) < tol
        assert abs(exc - exco) < 1e-10 * abs(exco)  # E_xc is FP64 in both modes
        g.close()


def test_c30h62_full_grid_builds():
    mol, P = _mol_and_guess()
    grid = (75, 302, 8)
    print("[case] C30H62 6-31G 75x302 FP64 dense", flush=True)
    g = G.MolecularGrid(mol, *grid)
    assert g.npts == 2083800
    J, XC, sc = _device_build(g, P)
    assert np.array_equal(J, J.T) and np.array_equal(XC, XC.T)
    assert abs(sc[2] - 242.0) < 1e-9
    assert abs(sc[3] - 2 * np.sum(P * J)) < 1e-10 * abs(sc[3])
    # the host-call sequence gives the same matrices as the fused device build
    Jh, XCh, exch, _ = _host_build(g, P)
    assert rel_err(Jh, J) < 1e-13 and rel_err(XCh, XC) < 1e-13 and abs(exch - sc[0]) < 1e-11 * abs(sc[0])
    g.close()
    print("[case] C30H62 75x302 FP64 screened 1e-10", flush=True)
    gs = G.MolecularGrid(mol, *grid, screen=1e-10)
    Js, XCs, ss = _device_build(gs, P)
    gs.close()
    assert rel_err(Js, J) < 1e-9 and rel_err(XCs, XC) < 1e-9 and abs(ss[0] - sc[0]) < 1e-8
    print("[case] C30H62 75x302 mixed", flush=True)
    gm = G.MolecularGrid(mol, *grid, precision=G.PREC_MIXED)
    Jm, XCm, sm = _device_build(gm, P)
    gm.close()
    print(f"  mixed: J {rel_err(Jm, J):.2e} XC {rel_err(XCm, XC):.2e}", flush=True)
    assert rel_err(Jm, J) < 1e-6 and rel_err(XCm, XC) < 1e-6
    assert np.array_equal(Jm, Jm.T) and np.array_equal(XCm, XCm.T)
    # E_xc and the pointwise E_J come from the same FP64 density and potential
    assert abs(sm[0] - sc[0]) < 1e-12 * abs(sc[0]) and abs(sm[3] - sc[3]) < 1e-12 * abs(sc[3])


def test_c30h62_full_grid_adjoint():
    mol, P = _mol_and_guess()
    grid = (75, 302, 8)
    print("[case] C30H62 75x302 adjoint", flush=True)
    g = G.MolecularGrid(mol, *grid)
    g.set_density(P)
    c = float(np.sum(mol.Z)) / g.correct_densities()
    XC, exc = g.calculate_exchange_correlation()
    dP_c, _ = g.xc_adjoint(hold_c=True, grad=False)
    assert np.abs(dP_c - 2 * c * XC).max() < 1e-12 * np.abs(XC).max()
    _, grad = g.xc_adjoint(hold_c=False, dEdP=False)
    g.close()
    assert np.abs(grad.sum(0)).max() < 1e-9 * np.abs(grad).max()
    h = 1e-4
    # a middle carbon along the chain and a hydrogen out of the plane
    for a, d in ((15, 0), (50, 2)):
        e = []
        for sgn in (1, -1):
            m2 = M.Molecule(mol.Z, mol.xyz.copy(), mol.basis, mol.name)
            m2.xyz[a, d] += sgn * h
            gd = G.MolecularGrid(m2, *grid)
            gd.set_density(P)
            gd.correct_densities()
            e.append(gd.calculate_exchange_correlation()[1])
            gd.close()
        fd = (e[0] - e[1]) / (2 * h)
        print(f"  atom {a} dir {d}: analytic {grad[a, d]:.10f} fd {fd:.10f}", flush=True)
        assert abs(fd - grad[a, d]) < 1e-5, (a, d, fd, grad[a, d])
