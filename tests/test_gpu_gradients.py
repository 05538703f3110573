"""Basis-function and density gradients on the GPU against the oracle
(src/gridpoint.cpp:62-108, src/rectangulargrid.cpp:34-95).

dg_get_amplitude_gradients / dg_get_density_gradients on the molecular grid,
and dg_density_at_points on the reference's rectangular grid, compared with
or_points_density at the same points (FP64: 1e-12 of the largest value).
"""
import numpy as np
import pytest

from dftcxx_amd import gridlib as G
from dftcxx_amd import molecule as M

from systems import random_density, rel_err, with_d_shell

pytestmark = pytest.mark.gpu


def _co_d():
    return with_d_shell(M.Molecule.from_angstrom(["C", "O"], [[0, 0, 0], [0, 0, 1.13]], "p631"))


@pytest.mark.parametrize("make", [lambda: M.h2o("sto3g"), _co_d], ids=["h2o", "co_d"])
def test_grid_gradients_vs_oracle(make):
    from oracle import oracle as O
    mol = make()
    print(f"gradients on {mol.name}: {mol.nbf()} functions", flush=True)
    g = G.MolecularGrid(mol, 10, 26, 2)
    P = random_density(mol.nbf(), 21)
    g.set_density(P)
    pts = g.get_positions()
    rho_o, grad_o, dphi_o = O.points_density(mol, pts, P)
    dphi = g.get_amplitude_gradients()
    assert rel_err(dphi, dphi_o) < 1e-12
    # before the correction: the gradient of rho = 2 phi^T P phi
    gr = g.get_density_gradients()
    assert rel_err(gr, grad_o) < 1e-12
    assert rel_err(g.get_densities(), rho_o) < 1e-12
    # after it: the same factor as the densities
    g.correct_densities()
    c = g.get_densities().sum() / rho_o.sum()
    gr2 = g.get_density_gradients()
    assert rel_err(gr2, c * grad_o) < 1e-12


def test_gradient_needs_a_density():
    g = G.MolecularGrid(M.h2o("sto3g"), 5, 6, 0)
    with pytest.raises(G.DgError, match="no density"):
        g.get_density_gradients()


@pytest.mark.parametrize("make", [lambda: M.h2o("p631"), _co_d], ids=["h2o", "co_d"])
def test_rectangular_grid_vs_oracle(make, tmp_path):
    from oracle import oracle as O
    mol = make()
    P = random_density(mol.nbf(), 4)
    rg = G.RectangularGrid(mol)
    rg.build_grid(5.0, 15)  # the reference's commented-out call (src/dft.cpp:493-503)
    pts_o = O.rectangular_points(5.0, 15)
    assert np.abs(rg.positions - pts_o).max() < 1e-15
    rg.set_density(P)
    rho_o, grad_o, _ = O.points_density(mol, rg.positions, P)
    assert rel_err(rg.densities, rho_o) < 1e-12
    assert rel_err(rg.gradients, grad_o) < 1e-12
    f = tmp_path / "data.dat"
    rg.write_gradient(str(f))
    rows = f.read_text().splitlines()
    assert len(rows) == 15 ** 3
    first = [float(v) for v in rows[0].split()]
    assert np.allclose(first[:3], rg.positions[0]) and np.allclose(first[3:], rg.gradients[0], atol=1e-8)


def test_points_in_chunks_and_empty():
    from oracle import oracle as O
    mol = M.h2o("sto3g")
    P = random_density(mol.nbf(), 8)
    # more points than one device chunk (65536), and a count not a multiple of 64
    pts = np.random.default_rng(2).uniform(-3, 3, (70001, 3))
    rho, grad = G.density_at_points(mol, pts, P)
    sel = np.r_[0:50, 65500:65600, 69950:70001]
    rho_o, grad_o, _ = O.points_density(mol, pts[sel], P)
    assert rel_err(rho[sel], rho_o) < 1e-12 and rel_err(grad[sel], grad_o) < 1e-12
    r0, g0 = G.density_at_points(mol, np.zeros((0, 3)), P)
    assert r0.shape == (0,) and g0.shape == (0, 3)


def test_density_gradients_after_forces():
    """The force terms form T = Phi P only on the tiles they read (nine tiles
    here, so the streamed T kernel with its per-block tile mask); a later
    dg_get_density_gradients forms all of T again and matches a fresh handle."""
    mol = M.alkane(10)
    P = random_density(mol.nbf(), 5)
    print(f"density gradients after forces on {mol.name}: {mol.nbf()} functions", flush=True)
    ref = G.MolecularGrid(mol, 20, 26, 4)
    ref.set_density(P)
    ref.correct_densities()
    want = ref.get_density_gradients()
    g = G.MolecularGrid(mol, 20, 26, 4)
    g.set_density(P)
    g.correct_densities()
    g.calculate_exchange_correlation()
    _, grad = g.xc_adjoint(hold_c=False)
    assert np.all(np.isfinite(grad))
    got = g.get_density_gradients()
    assert np.array_equal(got, want)
