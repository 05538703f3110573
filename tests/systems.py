"""Test molecules and helpers shared by the test modules."""
import numpy as np

from dftcxx_amd import molecule as M


def random_density(nbf, seed):
    """Seeded symmetric positive semi-definite half density matrix."""
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((nbf, max(1, nbf // 2))) * 0.3
    return C @ C.T


def core_guess(mol, S, T, V):
    """Half density matrix from the core Hamiltonian (src/dft.cpp:330-366, J = XC = 0)."""
    w, U = np.linalg.eigh(S)
    X = U / np.sqrt(w)
    e, Cc = np.linalg.eigh(X.T @ (T + V) @ X)
    C = X @ Cc
    nocc = mol.nelec // 2
    return C[:, :nocc] @ C[:, :nocc].T


def with_d_shell(mol):
    """Copy of mol whose heavy atoms carry an extra Cartesian d shell."""
    table = {z: list(v) for z, v in M.load_basis(mol.basis).items()}
    for z in set(int(v) for v in mol.Z):
        if z > 2:
            table[z] = table[z] + [(2, [0.8, 0.25], [0.6, 0.5])]
    return M.Molecule(mol.Z, mol.xyz, table, mol.name + "+d")


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
