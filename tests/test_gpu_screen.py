"""GPU checks of amplitude screening (dg_grid_params.screen): the screened
builds against the dense FP64 build of the same molecule and grid."""
import numpy as np
import pytest

from dftcxx_amd import gridlib as G
from dftcxx_amd import molecule as M

from systems import core_guess

pytestmark = pytest.mark.gpu

SCREEN = 1e-10


def _build(mol, grid, P, precision=G.PREC_FP64, screen=0.0):
    g = G.MolecularGrid(mol, *grid, precision=precision, screen=screen)
    g.set_density(P)
    g.correct_densities()
    J = g.calculate_hartree_potential()
    XC, exc = g.calculate_exchange_correlation()
    return J, XC, exc, g.get_densities()


@pytest.mark.parametrize("make", [lambda: M.benzene("p631"), lambda: M.alkane(10)], ids=["benzene", "c10h22"])
def test_screened_fp64_build_matches_dense(make):
    mol = make()
    grid = (75, 302, 8)
    P = core_guess(mol, *G.one_electron(mol))
    print(f"[case] screen {mol.name}", flush=True)
    J0, X0, e0, r0 = _build(mol, grid, P)
    J1, X1, e1, r1 = _build(mol, grid, P, screen=SCREEN)
    assert np.abs(r1 - r0).max() < 1e-9 * r0.max()
    assert np.abs(J1 - J0).max() < 1e-9 * np.abs(J0).max()
    assert np.abs(X1 - X0).max() < 1e-9 * np.abs(X0).max()
    assert abs(e1 - e0) < 1e-8


def test_screened_scf_energies():
    mol = M.alkane(4)
    grid = (50, 110, 6)
    print("[case] screened SCF butane", flush=True)
    # converged far below the tolerances, so that only the build's errors remain
    kw = dict(threshold=1e-10, max_iter=400)
    _, ref, _ = G.scf_run(mol, *grid, **kw)
    _, scr, _ = G.scf_run(mol, *grid, screen=SCREEN, **kw)
    _, mix, _ = G.scf_run(mol, *grid, precision=G.PREC_MIXED, screen=SCREEN, **kw)
    _, mx0, _ = G.scf_run(mol, *grid, precision=G.PREC_MIXED, **kw)
    print("  iterations", ref["iterations"], scr["iterations"], mix["iterations"], mx0["iterations"])
    print("  errors", scr["energy"] - ref["energy"], mix["energy"] - ref["energy"], mx0["energy"] - ref["energy"])
    assert max(r["iterations"] for r in (ref, scr)) < 400
    assert abs(scr["energy"] - ref["energy"]) < 1e-8
    assert abs(mx0["energy"] - ref["energy"]) < 1e-6
    assert abs(mix["energy"] - ref["energy"]) < 1e-6


def test_screen_must_be_non_negative():
    with pytest.raises(G.DgError):
        G.MolecularGrid(M.h2o(), 10, 26, 4, screen=-1.0)
