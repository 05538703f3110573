"""The reference-shaped C++ facade (include/dftcxx/molecular_grid.hpp) on the GPU.

tests/facade/libfacade_probe.so includes the header-only facade and makes the
calls DFT makes on MolecularGrid (src/dft.cpp:57-61, 362-365, 395-397, 483):
MolecularGrid(mol), set_grid_parameters(50, 7, 8) with the reference's
Lebedev enum index (7 -> 110 points), create_grid, set_density,
correct_densities, calculate_hartree_potential, the XC assembly and
calculate_density.  Its J and XC are compared with the ctypes binding and the
oracle on H2O / STO-3G / 50 x 110, the configuration the reference runs.
"""
import ctypes as C
import os

import numpy as np
import pytest

from dftcxx_amd import gridlib as G
from dftcxx_amd import molecule as M

from systems import core_guess, rel_err

pytestmark = pytest.mark.gpu

PROBE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "facade", "libfacade_probe.so")


def test_cpp_facade_build_matches_ctypes_and_oracle():
    from oracle import oracle as O
    assert os.path.exists(PROBE), "build tests/facade first (__graft_entry__.build())"
    G.lib()  # the facade probe binds to the same libdftgrid.so
    L = C.CDLL(PROBE)
    L.facade_probe_error.restype = C.c_char_p
    mol = M.h2o("sto3g")
    P = core_guess(mol, *G.one_electron(mol))
    s = mol.shells()
    keep = [np.ascontiguousarray(mol.Z, np.int32), np.ascontiguousarray(mol.xyz, np.float64)]
    keep += [np.ascontiguousarray(s[k], np.int32) for k in ("atom", "L", "nprim", "offset")]
    keep += [np.ascontiguousarray(s[k], np.float64) for k in ("alpha", "coef")]
    n = mol.nbf()
    J, XC, sc = np.empty(n * n), np.empty(n * n), np.empty(3)
    Pc = np.ascontiguousarray(P.T)
    print("[case] facade h2o 50x110 (lebedev_order 7)", flush=True)
    rc = L.facade_probe(mol.natoms, G._i(keep[0]), G._d(keep[1]), len(keep[2]), G._i(keep[2]), G._i(keep[3]),
                        G._i(keep[4]), G._i(keep[5]), G._d(keep[6]), G._d(keep[7]), len(keep[6]),
                        50, 7, 8, G._d(Pc), G._d(J), G._d(XC), G._d(sc))
    assert rc == 0, L.facade_probe_error().decode()
    J = J.reshape(n, n).T
    XC = XC.reshape(n, n).T
    g = G.MolecularGrid(mol)
    g.set_grid_parameters(50, 7, 8)
    g.create_grid()
    g.set_density(P)
    nb = g.correct_densities()
    Jg = g.calculate_hartree_potential()
    XCg, excg = g.calculate_exchange_correlation()
    assert np.array_equal(J, Jg) and np.array_equal(XC, XCg)
    assert sc[0] == excg and sc[1] == nb and abs(sc[2] - mol.nelec) < 1e-12
    o = O.Grid(mol, 50, 110, 8)
    o.set_density(P)
    o.correct_densities()
    assert rel_err(J, o.hartree()) < 1e-10
    XCo, exco = o.xc()
    assert rel_err(XC, XCo) < 1e-10 and abs(sc[0] - exco) < 1e-10
