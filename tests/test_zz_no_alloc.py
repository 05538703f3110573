"""dg_density_at_points_noalloc (work buffers in the library's static device
array, no allocation) against dg_density_at_points (the allocating entry
point it is a variant of), through the C ABI: two calls in a row, the second
with fewer points, so that the second runs on an array holding the first
call's contents; then a basis too large for the array is refused."""
import numpy as np
import pytest

from dftcxx_amd import gridlib as G
from dftcxx_amd import molecule as M

from systems import random_density, rel_err

pytestmark = pytest.mark.gpu


def _pair(mol, pts, P):
    r_a, g_a = G.density_at_points(mol, pts, P)
    r_n, g_n = G.density_at_points(mol, pts, P, noalloc=True)
    return r_a, g_a, r_n, g_n


def test_noalloc_matches_allocating_entry_point():
    rng = np.random.default_rng(17)
    big = M.alkane(10)  # 134 functions: several passes through the scratch array
    P1 = random_density(big.nbf(), 3)
    pts1 = rng.uniform(-4.0, 14.0, (40000, 3))
    print(f"no-alloc call 1: {big.nbf()} functions, {len(pts1)} points", flush=True)
    r_a, g_a, r_n, g_n = _pair(big, pts1, P1)
    assert np.array_equal(r_n, r_a) and np.array_equal(g_n, g_a)
    small = M.h2o("sto3g")
    P2 = random_density(small.nbf(), 5)
    pts2 = rng.uniform(-2.0, 2.0, (999, 3))
    print(f"no-alloc call 2: {small.nbf()} functions, {len(pts2)} points", flush=True)
    r_a, g_a, r_n, g_n = _pair(small, pts2, P2)
    assert np.array_equal(r_n, r_a) and np.array_equal(g_n, g_a)
    from oracle import oracle as O
    rho_o, grad_o, _ = O.points_density(small, pts2, P2)
    assert rel_err(r_n, rho_o) < 1e-12 and rel_err(g_n, grad_o) < 1e-12


def test_noalloc_refuses_oversized_basis():
    table = {1: [(2, [1.0 + 0.1 * k], [1.0]) for k in range(90)]}  # 540 functions on one atom
    mol = M.Molecule([1], [[0.0, 0.0, 0.0]], table, "too many")
    with pytest.raises(G.DgError, match="more than 512"):
        G.density_at_points(mol, np.zeros((4, 3)), np.eye(mol.nbf()), noalloc=True)
