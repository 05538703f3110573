"""GPU density for bases of more than 112 functions (the streamed-operand
k_density path: 8 or more 16-function tiles, fragments double-buffered in LDS)
against rho = 2 diag(Phi P Phi^T) evaluated in FP64 with numpy from the
library's own amplitudes (which test_gpu_parity.py checks against the oracle),
dense and screened.  Small grids keep the numpy reference quick."""
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
