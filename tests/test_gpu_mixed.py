"""Reduced-precision (DG_PREC_MIXED) grid build against the FP64 build.

Mixed mode keeps rho, the functional, the Poisson solve, the Hartree
potential and every energy in FP64; the Fock contractions J and XC run on the
FP32 16x16x4 MFMA (FP32 amplitudes, FP32 sums added into FP64 every 64 points
per wave).  Matrix errors enter the SCF energy at second order only (E_J is
evaluated pointwise in FP64).  Tolerances: J and XC within 1e-6 of their
scale; total SCF energies within 1e-6 Ha of the FP64 build.
"""
import numpy as np
import pytest

from dftcxx_amd import gridlib as G
from dftcxx_amd import molecule as M

from systems import core_guess, rel_err

pytestmark = pytest.mark.gpu

CASES = [
    ("h2o_sto3g_50x110", lambda: M.h2o("sto3g"), 50, 110, 8),
    ("ch4_631g_30x50", lambda: M.ch4("p631"), 30, 50, 6),
    ("benzene_sto3g_20x50", lambda: M.benzene("sto3g"), 20, 50, 6),
]


def _build(g, P):
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


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_mixed_build_matches_fp64(case):
    name, make, nrad, nang, lmax = case
    print(f"[case] {name}", flush=True)
    mol = make()
    S, T, V = G.one_electron(mol)
    P = core_guess(mol, S, T, V)
    g64 = G.MolecularGrid(mol, nrad, nang, lmax, precision=G.PREC_FP64)
    gmx = G.MolecularGrid(mol, nrad, nang, lmax, precision=G.PREC_MIXED)
    J64, X64, s64 = _build(g64, P)
    Jmx, Xmx, smx = _build(gmx, P)
    # identity 2 tr(P J) = sum w V rho / (2c) in FP64
    assert abs(s64[3] - 2 * np.trace(P @ J64)) < 1e-10 * abs(s64[3])
    assert rel_err(Jmx, J64) < 1e-6
    assert rel_err(Xmx, X64) < 1e-6
    assert np.abs(Jmx - Jmx.T).max() == 0.0 and np.abs(Xmx - Xmx.T).max() == 0.0
    assert abs(smx[0] - s64[0]) < 1e-12 * abs(s64[0])      # E_xc: same FP64 path
    assert abs(smx[3] - s64[3]) < 1e-12 * abs(s64[3])      # E_J pointwise: the same FP64 rho and V
    assert abs(smx[2] - mol.nelec) < 1e-10


@pytest.mark.parametrize("make,nrad,nang", [(lambda: M.h2o("sto3g"), 50, 110),
                                            (lambda: M.ch4("p631"), 40, 110)], ids=["h2o", "ch4"])
def test_mixed_scf_energy_within_1e6(make, nrad, nang):
    mol = make()
    e64, r64, P64 = G.scf_run(mol, nrad, nang, 8, threshold=1e-8, precision=G.PREC_FP64)
    emx, rmx, Pmx = G.scf_run(mol, nrad, nang, 8, threshold=1e-8, precision=G.PREC_MIXED)
    print(f"E64 {r64['energy']:.10f} Emixed {rmx['energy']:.10f} diff {rmx['energy'] - r64['energy']:.2e}")
    assert abs(rmx["energy"] - r64["energy"]) < 1e-6
    assert np.abs(Pmx - P64).max() < 1e-4


def test_fp32_mfma_layout():
    # FP32 16x16x4 MFMA: A[i][k] on lane i + 16k, B[k][j] on lane j + 16k,
    # D[4g + v][c] in register v of lane c + 16g (the layout k_contract_f32 assumes)
    rng = np.random.default_rng(4)
    a = rng.integers(-8, 8, 64).astype(np.float32)
    b = rng.integers(-8, 8, 64).astype(np.float32)
    A = np.array([[a[i + 16 * k] for k in range(4)] for i in range(16)], dtype=np.float64)
    B = np.array([[b[j + 16 * k] for j in range(16)] for k in range(4)], dtype=np.float64)
    D = A @ B
    d = G.mfma_probe_f32(a, b)
    for lane in range(64):
        c, g = lane % 16, lane // 16
        for v in range(4):
            assert d[v, lane] == D[4 * g + v, c], (lane, v)


@pytest.mark.parametrize("make,grid", [(lambda: M.h2o("sto3g"), (50, 110, 8)),
                                       (lambda: M.ch4("p631"), (30, 50, 6))], ids=["h2o", "ch4_631g"])
def test_mixed_scf_against_oracle(make, grid):
    # the reduced-precision SCF energy stated against the FP64 oracle's SCF
    from oracle import oracle as O
    mol = make()
    print(f"[case] mixed SCF {mol.name} {grid} vs oracle", flush=True)
    eo, to, Po = O.Grid(mol, *grid).scf(threshold=1e-8, max_iter=200)
    em, rm, Pm = G.scf_run(mol, *grid, threshold=1e-8, max_iter=200, precision=G.PREC_MIXED)
    print(f"  oracle {to['total']:.10f} mixed {rm['energy']:.10f} diff {rm['energy'] - to['total']:.2e}")
    assert rm["converged"] == 1
    assert abs(rm["energy"] - to["total"]) < 1e-6
    assert abs(rm["e_xc"] - to["xc"]) < 1e-6
    assert np.abs(Pm - Po).max() < 1e-4
