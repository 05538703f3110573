"""GPU density for bases of more than 112 functions (8 or more 16-function
tiles: k_density_ring without screening, the P fragments streamed through an LDS
ring; k_density with screening) against rho = 2 diag(Phi P Phi^T) evaluated in
FP64 with numpy from the library's own amplitudes (which test_gpu_parity.py
checks against the oracle), dense and screened.  Small grids keep the numpy
reference quick; the determinism checks run on full-size grids, where many
work-groups share each CU."""
import numpy as np
import pytest

from dftcxx_amd import gridlib as G
from dftcxx_amd import molecule as M

from systems import core_guess, random_density

pytestmark = pytest.mark.gpu


def _reference_density(g, P):
    phi = g.get_amplitudes()
    return 2.0 * np.einsum("pi,ij,pj->p", phi, P, phi, optimize=True)


@pytest.mark.parametrize("n_carbon,grid", [(10, (20, 26, 4)), (30, (10, 26, 2))], ids=["c10h22", "c30h62"])
@pytest.mark.parametrize("screen", [0.0, 1e-10])
def test_large_basis_density(n_carbon, grid, screen):
    mol = M.alkane(n_carbon)
    assert mol.nbf() > 112
    # a core guess plus a symmetric random part, so that every tile pair carries weight
    P = core_guess(mol, *G.one_electron(mol)) + 0.01 * random_density(mol.nbf(), seed=3)
    print(f"[case] density {mol.name} nbf={mol.nbf()} grid={grid} screen={screen}", flush=True)
    g = G.MolecularGrid(mol, *grid, screen=screen)
    g.set_density(P)
    rho = g.get_densities()
    ref = _reference_density(g, P)
    scale = np.abs(ref).max()
    tol = 1e-12 if screen == 0.0 else 1e-9
    assert np.abs(rho - ref).max() < tol * scale


def test_large_basis_density_is_deterministic():
    mol = M.alkane(10)
    P = core_guess(mol, *G.one_electron(mol))
    g = G.MolecularGrid(mol, 20, 26, 4)
    g.set_density(P)
    r1 = g.get_densities().copy()
    g.set_density(P)
    assert np.array_equal(g.get_densities(), r1)


@pytest.mark.parametrize("n_carbon,grid", [(10, (75, 302, 8)), (30, (40, 110, 4))], ids=["c10h22", "c30h62"])
def test_ring_kernels_are_deterministic_across_handles(n_carbon, grid):
    # the streamed-fragment density and T = Phi P (the forces) on several handles
    # and repeats: every result bitwise equal (a copy into a ring slot that another
    # wave was still reading would show up here as a difference)
    mol = M.alkane(n_carbon)
    P = core_guess(mol, *G.one_electron(mol))
    print(f"[case] ring determinism {mol.name} grid={grid}", flush=True)
    rhos, grads = [], []
    for _ in range(3):
        g = G.MolecularGrid(mol, *grid)
        for _ in range(2):
            g.set_density(P)
            rhos.append(g.get_densities().copy())
        g.correct_densities()
        g.calculate_exchange_correlation()
        grads.append(g.xc_adjoint(hold_c=True, dEdP=False)[1])
        g.close()
    for r in rhos[1:]:
        assert np.array_equal(r, rhos[0])
    for gr in grads[1:]:
        assert np.array_equal(gr, grads[0])
