"""The density-matrix step on the GPU (dg_scf_params.density_step = DG_STEP_DEVICE)
against the host eigen-solver step and the oracle SCF.

The device step forms the occupied projector by SP2 purification instead of
diagonalising; with a HOMO-LUMO gap it is the same P up to rounding, so the
energy sequences agree to well below the 1e-8 Ha of the FP64 parity bar.
Without a gap (a carbon atom: three degenerate 2p orbitals, one pair occupied;
the core-Hamiltonian guess of benzene) the projection cannot converge and the
step falls back to the host eigen-solver.
"""
import numpy as np
import pytest

from dftcxx_amd import gridlib as G
from dftcxx_amd import molecule as M

pytestmark = pytest.mark.gpu


def test_h2o_device_step_vs_host_and_oracle():
    from oracle import oracle as O
    mol = M.h2o("sto3g")
    en_h, r_h, P_h = G.scf_run(mol, 50, 110, 8, density_step=G.STEP_HOST)
    en_d, r_d, P_d = G.scf_run(mol, 50, 110, 8, density_step=G.STEP_DEVICE)
    assert r_d["device_step"] == 1 and r_h["device_step"] == 0
    assert r_d["host_fallbacks"] == 0
    assert len(en_d) == len(en_h) and r_d["converged"] == 1
    assert np.abs(en_d - en_h).max() < 1e-10
    assert np.abs(P_d - P_h).max() < 1e-9
    og = O.Grid(mol, 50, 110, 8)
    en_o, terms_o, P_o = og.scf(threshold=1e-6)
    assert len(en_d) == len(en_o)
    assert np.abs(en_d - en_o).max() < 1e-8
    assert abs(r_d["e_j"] - terms_o["J"]) < 1e-8 and abs(r_d["e_xc"] - terms_o["xc"]) < 1e-8


def test_device_step_vs_host_benzene():
    # the core-Hamiltonian guess of benzene has a degenerate HOMO/LUMO pair, so its
    # step falls back to the host (the same choice within the degenerate space as the
    # host path); every later step has a gap and runs on the device
    mol = M.benzene()
    en_h, r_h, P_h = G.scf_run(mol, 30, 110, 6, max_iter=8, threshold=0.0, density_step=G.STEP_HOST)
    en_d, r_d, P_d = G.scf_run(mol, 30, 110, 6, max_iter=8, threshold=0.0, density_step=G.STEP_DEVICE)
    assert 1 <= r_d["host_fallbacks"] < 8
    assert len(en_d) == len(en_h) == 8
    assert np.abs(en_d - en_h).max() < 1e-9
    assert np.abs(P_d - P_h).max() < 1e-8
    assert abs(r_d["nelec"] - mol.nelec) < 1e-9


def test_device_step_vs_host_c10h22():
    # 134 functions (the multi-work-group matrix products); the reference's SCF does
    # not settle for this chain (the energy wanders by tens of Ha), so the sequences are
    # compared over the first iterations, before rounding differences are amplified
    mol = M.alkane(10)
    en_h, r_h, _ = G.scf_run(mol, 20, 50, 4, max_iter=4, threshold=0.0, density_step=G.STEP_HOST)
    en_d, r_d, _ = G.scf_run(mol, 20, 50, 4, max_iter=4, threshold=0.0, density_step=G.STEP_DEVICE)
    assert len(en_d) == len(en_h) == 4
    assert np.abs(en_d - en_h).max() < 1e-8


def test_device_step_c30h62_first_iterations():
    # 394 functions; the first steps of the headline molecule against the host path
    # (its core-Hamiltonian guess has a HOMO-LUMO gap of 5e-8 Ha, so both P carry
    # rounding differences amplified by about 1/gap)
    mol = M.alkane(30)
    print("device step on C30H62", flush=True)
    en_h, _, _ = G.scf_run(mol, 10, 26, 2, max_iter=3, threshold=0.0, density_step=G.STEP_HOST)
    en_d, r_d, _ = G.scf_run(mol, 10, 26, 2, max_iter=3, threshold=0.0, density_step=G.STEP_DEVICE)
    assert np.abs(en_d - en_h).max() < 1e-6


def test_device_step_mixed_precision():
    # mixed mode forms G = 2J + XC in one FP32 contraction on the device path and J, XC
    # in two on the host path: both within the 1e-6 Ha bar of the FP64 SCF
    mol = M.alkane(4)
    en_f, _, _ = G.scf_run(mol, 20, 50, 4, density_step=G.STEP_HOST)
    en_h, r_h, _ = G.scf_run(mol, 20, 50, 4, precision=G.PREC_MIXED, density_step=G.STEP_HOST)
    en_d, r_d, _ = G.scf_run(mol, 20, 50, 4, precision=G.PREC_MIXED, density_step=G.STEP_DEVICE)
    assert r_d["converged"] == 1 and r_h["converged"] == 1
    assert abs(r_d["energy"] - en_f[-1]) < 1e-6 and abs(r_h["energy"] - en_f[-1]) < 1e-6


def test_no_gap_falls_back_to_host():
    # a carbon atom: three degenerate 2p orbitals, one pair occupied; no step has a gap
    mol = M.Molecule([6], [[0.0, 0.0, 0.0]], "sto3g", "C atom")
    en_h, r_h, _ = G.scf_run(mol, 20, 50, 4, max_iter=6, threshold=0.0, density_step=G.STEP_HOST)
    en_d, r_d, _ = G.scf_run(mol, 20, 50, 4, max_iter=6, threshold=0.0, density_step=G.STEP_DEVICE)
    # at least the set-up step (once the SCF has occupied one 2p orbital, the grid's
    # Lebedev directions split the level and later steps may have a gap)
    assert r_d["host_fallbacks"] >= 1
    # the host's choice within the degenerate space depends on F at rounding level,
    # and F differs there between the two paths (one contraction or two)
    assert np.abs(en_d - en_h).max() < 1e-5


@pytest.mark.parametrize("make,prec", [(M.benzene, G.PREC_FP64), (lambda: M.alkane(10), G.PREC_FP64),
                                       (lambda: M.alkane(10), G.PREC_MIXED)], ids=["benzene", "c10h22", "c10h22_mixed"])
def test_fock_build_is_2j_plus_xc(make, prec):
    import torch
    from systems import random_density
    mol = make()
    g = G.MolecularGrid(mol, 20, 50, 4, precision=prec)
    n = mol.nbf()
    dev = torch.device("cuda:0")
    P = torch.tensor(random_density(n, 7).T.copy(), dtype=torch.float64, device=dev)
    J = torch.zeros((n, n), dtype=torch.float64, device=dev)
    XC = torch.zeros_like(J)
    Gm = torch.zeros_like(J)
    s1 = torch.zeros(4, dtype=torch.float64, device=dev)
    s2 = torch.zeros_like(s1)
    torch.cuda.synchronize()
    g.build_device(P.data_ptr(), J.data_ptr(), XC.data_ptr(), s1.data_ptr())
    g.build_fock_device(P.data_ptr(), Gm.data_ptr(), s2.data_ptr())
    g.close()
    torch.cuda.synchronize()
    ref = (2 * J + XC).cpu().numpy()
    got = Gm.cpu().numpy()
    # mixed: one FP32 contraction against two, each FP32 sum flushed to FP64 every 64
    # points, so the two differ at a few FP32 ulps of the larger terms
    tol = 1e-12 if prec == G.PREC_FP64 else 5e-7
    assert np.abs(got - ref).max() < tol * np.abs(ref).max()
    assert np.array_equal(s1.cpu().numpy(), s2.cpu().numpy())
