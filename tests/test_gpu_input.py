"""The reference's own input files run end to end: dg_input_read -> dg_scf_run on
the GPU, against the oracle SCF on the same parsed molecule and grid.

h2o.in is the reference's water input (6-31G, bohr, the "medium" grid of
15 radial x 110 Lebedev points, lmax 8); lih.in its lithium hydride (STO-3G).
"""
import os

import numpy as np
import pytest

from dftcxx_amd import gridlib as G

pytestmark = pytest.mark.gpu
MOLS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "molecules")


@pytest.mark.parametrize("name", ["h2o", "lih"])
def test_scf_from_reference_input(name):
    from oracle import oracle as O
    f = G.read_input(os.path.join(MOLS, name + ".in"))
    print(f"input {name}: {f.natoms} atoms, {f.nbf()} functions, grid {f.radial_points}x{f.angular_points} "
          f"lmax {f.lmax}", flush=True)
    og = O.Grid(f, f.radial_points, f.angular_points, f.lmax)
    en_o, terms_o, P_o = og.scf(threshold=1e-6)
    en_g, res, P_g = G.scf_run(f, f.radial_points, f.angular_points, f.lmax, threshold=1e-6)
    assert res["converged"] == 1
    assert len(en_g) == len(en_o)
    assert np.abs(en_g - en_o).max() < 1e-8
    assert abs(res["energy"] - terms_o["total"]) < 1e-8
    assert np.abs(P_g - P_o).max() < 1e-7
