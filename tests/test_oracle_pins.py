"""Pins of the oracle (FP64 CPU restatement) against mathematics and libraries.

The reference ships no tests, golden vectors or runnable binary (its build
needs Eigen, Boost and TCLAP, absent here), so each oracle function is pinned
by an independent check chosen so that a plausible mistake -- a dropped term,
a wrong sign or index, a transposed operand -- fails it:
closed forms, invariants, exact rational arithmetic, SciPy routines, and
convergence in the grid size.
"""
import math
import os
from fractions import Fraction

import numpy as np
import pytest
from scipy import integrate, special
from scipy.interpolate import CubicSpline

from dftcxx_amd import lebedev
from dftcxx_amd import molecule as M

from systems import core_guess, random_density

REF_QUAD = "/root/reference/src/quadrature.h"


# --- a1: Lebedev tables --------------------------------------------------------

def test_lebedev_table_regenerates_from_scipy():
    from dftcxx_amd.data.make_lebedev import ORDERS, lebedev_table
    for n, deg in ORDERS:
        assert np.array_equal(lebedev_table(deg), lebedev.rule(n)), n


@pytest.mark.skipif(not os.path.exists(REF_QUAD), reason="reference tree not mounted")
def test_lebedev_matches_reference_tables():
    import re
    txt = open(REF_QUAD).read()
    rows = re.findall(r"\{\s*([-\d.e]+),\s*([-\d.e]+),\s*([-\d.e]+),\s*([-\d.e]+)\}", txt)
    table = np.array(rows, dtype=float)
    off = 0
    for n in lebedev.REFERENCE_ORDERS:
        assert np.array_equal(table[off:off + n], lebedev.rule(n)), n
        off += n


@pytest.mark.parametrize("n,deg", [(110, 17), (194, 23), (302, 29)])
def test_lebedev_integrates_harmonic_products(oracle_lib, n, deg):
    t = lebedev.rule(n)
    assert abs(t[:, 3].sum() - 1.0) < 1e-10
    assert np.abs(np.linalg.norm(t[:, :3], axis=1) - 1.0).max() < 2e-12
    lmax = deg // 2
    Y = np.array([[oracle_lib.ylm_xyz(l, m, *p) for l in range(lmax + 1) for m in range(-l, l + 1)]
                  for p in t[:, :3]])
    G = (Y * (4 * np.pi * t[:, 3:4])).T @ Y
    # 12-decimal weights limit the exactness to ~1e-10
    assert np.abs(G - np.eye(len(G))).max() < 2e-10


# --- a2: radial grid -------------------------------------------------------------

def _single_atom(nrad, nang=110, basis="sto3g"):
    from oracle import oracle as O
    mol = M.Molecule([1], [[0.3, -0.2, 0.1]], basis)
    return mol, O.Grid(mol, nrad, nang, 4)


def test_radial_nodes_closed_form(oracle_lib):
    for N in (10, 50, 75):
        _, g = _single_atom(N, 6)
        x = np.cos(np.arange(1, N + 1) * np.pi / (N + 1))
        assert np.abs(g.radial() - (1 + x) / (1 - x)).max() < 1e-12 * ((1 + x) / (1 - x)).max()


def test_radial_quadrature_known_answers(oracle_lib):
    # single centre, Becke weight 1: sum w f(r) -> int f d^3r
    errs = []
    for N in (25, 50, 100):
        mol, g = _single_atom(N)
        xyz, w, wb = g.points()
        assert np.all(wb == 1.0)
        r2 = ((xyz - mol.xyz[0]) ** 2).sum(1)
        errs.append(abs((w * np.exp(-r2)).sum() - np.pi ** 1.5))
        # int r^2 exp(-2 r^2) d^3r = 4 pi * Gamma(5/2) / (2 * 2^{5/2})
        exact = 4 * np.pi * math.gamma(2.5) / (2 * 2 ** 2.5)
        assert abs((w * r2 * np.exp(-2 * r2)).sum() - exact) < 1e-3 * exact
    assert errs[-1] < 1e-9 and errs[0] > errs[-1]


# --- a3: Becke partition -----------------------------------------------------------

def _s_exact(mu):
    for _ in range(3):
        mu = Fraction(3, 2) * mu - Fraction(1, 2) * mu ** 3
    return Fraction(1, 2) * (1 - mu)


def test_becke_cutoff_exact_rational(oracle_lib):
    for mu in (Fraction(0), Fraction(1), Fraction(-1), Fraction(1, 3), Fraction(-7, 10),
               Fraction(123, 1000)):
        assert abs(oracle_lib.becke_cutoff(float(mu)) - float(_s_exact(mu))) < 1e-15
    for mu in np.linspace(-1, 1, 41):
        assert abs(oracle_lib.becke_cutoff(mu) + oracle_lib.becke_cutoff(-mu) - 1.0) < 1e-15


def test_becke_partition_of_unity(oracle_lib):
    # integrals of a function centred anywhere are recovered by the fuzzy cells
    from oracle import oracle as O
    mol = M.h2o("sto3g")
    g = O.Grid(mol, 60, 194, 4)
    xyz, w, wb = g.points()
    assert np.all((wb >= 0) & (wb <= 1))
    for centre in (mol.xyz[0], mol.xyz[1], 0.5 * (mol.xyz[1] + mol.xyz[2])):
        r2 = ((xyz - centre) ** 2).sum(1)
        assert abs((w * np.exp(-1.3 * r2)).sum() - (np.pi / 1.3) ** 1.5) < 1e-5


def test_becke_homonuclear_symmetry(oracle_lib):
    # H2: the two cells are mirror images, so their populations are equal
    from oracle import oracle as O
    mol = M.Molecule([1, 1], [[0, 0, -0.7], [0, 0, 0.7]], "sto3g")
    g = O.Grid(mol, 30, 50, 2)
    P = np.array([[0.3, 0.3], [0.3, 0.3]])
    g.set_density(P)
    q = g.cell_densities()
    assert abs(q[0] - q[1]) < 1e-13
    xyz, w, wb = g.points()
    n = g.npts // 2
    # the mirror z -> -z maps cell 0 onto cell 1 (Lebedev point j onto its z-flipped partner)
    assert np.allclose(np.sort(wb[:n]), np.sort(wb[n:]), rtol=0, atol=1e-14)


# --- a4: basis amplitudes -----------------------------------------------------------

def test_gto_norm_closed_form(oracle_lib):
    # N^2 * int x^2l y^2m z^2n exp(-2 a r^2) = (pi / 3.14159265359)^{3/2}
    for a in (0.17, 1.0, 38.4):
        for l, m, n in ((0, 0, 0), (1, 0, 0), (0, 1, 0), (2, 0, 0), (1, 1, 0), (0, 1, 1)):
            L = l + m + n
            df = lambda k: float(np.prod(np.arange(2 * k - 1, 0, -2))) if k > 0 else 1.0
            integral = df(l) * df(m) * df(n) * np.pi ** 1.5 / (2 ** L * (2 * a) ** (L + 1.5))
            N = oracle_lib.gto_norm(a, l, m, n)
            assert abs(N * N * integral - (np.pi / 3.14159265359) ** 1.5) < 1e-13


def test_amplitudes_point_values(oracle_lib):
    from oracle import oracle as O
    mol = M.co("p631")
    g = O.Grid(mol, 12, 14, 2)
    phi = g.amplitudes()
    xyz, _, _ = g.points()
    sh = mol.shells()
    rng = np.random.default_rng(3)
    for p in rng.integers(0, g.npts, 20):
        col = 0
        for s in range(len(sh["L"])):
            A = mol.xyz[sh["atom"][s]]
            d = xyz[p] - A
            r2 = d @ d
            comps = {0: [(0, 0, 0)], 1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)]}[int(sh["L"][s])]
            for lmn in comps:
                val = 0.0
                for k in range(sh["nprim"][s]):
                    a = sh["alpha"][sh["offset"][s] + k]
                    c = sh["coef"][sh["offset"][s] + k]
                    val += c * oracle_lib.gto_norm(a, *lmn) * np.prod(d ** np.array(lmn)) * np.exp(-a * r2)
                assert abs(phi[p, col] - val) < 1e-14 * max(1.0, abs(val))
                col += 1


def test_amplitude_overlap_matches_analytic(oracle_lib):
    from oracle import oracle as O
    mol = M.h2o("sto3g")
    S, _, _ = O.one_electron(mol)
    errs = []
    for nrad, nang in ((30, 50), (60, 146), (100, 302)):
        g = O.Grid(mol, nrad, nang, 2)
        phi = g.amplitudes()
        _, w, _ = g.points()
        errs.append(np.abs(phi.T @ (w[:, None] * phi) - S).max())
    assert errs[-1] < 1e-6 and errs[0] > errs[-1]


# --- a5/a6: density and renormalisation ----------------------------------------------

def test_density_trace_and_correction(oracle_lib):
    from oracle import oracle as O
    mol = M.h2o("sto3g")
    S, T, V = O.one_electron(mol)
    P = core_guess(mol, S, T, V)
    g = O.Grid(mol, 100, 302, 2)
    g.set_density(P)
    rho = g.densities()
    assert rho.min() >= 0.0
    # int rho = 2 tr(P S) up to grid error (= nelec for the SCF density)
    assert abs(g.calculate_density() - 2 * np.trace(P @ S)) < 1e-6
    g.correct_densities()
    assert abs(g.calculate_density() - mol.nelec) < 1e-12


# --- a7/a8: LDA -----------------------------------------------------------------------

def test_lda_closed_forms_and_derivatives(oracle_lib):
    rho = np.logspace(-8, 3, 23)
    ex, ec, vxa, vca = oracle_lib.lda(rho)
    # Dirac exchange of the unpolarised gas
    assert np.allclose(ex, -0.75 * (3 / np.pi) ** (1 / 3) * rho ** (4 / 3), rtol=1e-13, atol=0)
    # v = d(e)/d(rho) (closed shell: v_a = dE/drho_a = dE/drho)
    h = 1e-6 * rho
    exp_, ecp, _, _ = oracle_lib.lda(rho + h)
    exm, ecm, _, _ = oracle_lib.lda(rho - h)
    assert np.allclose(vxa, (exp_ - exm) / (2 * h), rtol=1e-7)
    assert np.allclose(vca, (ecp - ecm) / (2 * h), rtol=1e-6, atol=1e-12)


def test_vwn_high_density_limit(oracle_lib):
    # eps_c -> A ln(r_s) + const as r_s -> 0 with A = (1 - ln 2)/pi^2 (paramagnetic)
    A = (1 - math.log(2)) / math.pi ** 2
    assert abs(A - 0.0310907) < 5e-8
    rs = np.array([1e-5, 2e-5])
    rho = 3 / (4 * np.pi * rs ** 3)
    _, ec, _, _ = oracle_lib.lda(rho)
    eps = ec / rho
    slope = (eps[1] - eps[0]) / math.log(2)
    assert abs(slope - 0.0310907) < 2e-4


def test_lda_density_thresholds(oracle_lib):
    # a spin channel below 1e-10 contributes nothing (src/functionals.cpp:30,42,82)
    rho = np.array([1.999e-10, 2.0e-10, 2.001e-10, 0.0])
    ex, ec, vxa, vca = oracle_lib.lda(rho)
    assert ex[0] == 0.0 and vxa[0] == 0.0 and ex[3] == 0.0 and ec[3] == 0.0
    assert ex[1] != 0.0 and ex[2] != 0.0
    assert ec[0] != 0.0  # correlation uses the total density


# --- a10: spherical harmonics ----------------------------------------------------------

def test_real_harmonics_match_scipy(oracle_lib):
    rng = np.random.default_rng(5)
    for _ in range(20):
        v = rng.standard_normal(3)
        r = np.linalg.norm(v)
        theta, phi = math.acos(v[2] / r), math.atan2(v[1], v[0])
        for l in range(9):
            for m in range(-l, l + 1):
                Yc = special.sph_harm_y(l, abs(m), theta, phi)
                ref = Yc.real if m == 0 else math.sqrt(2) * (Yc.real if m > 0 else Yc.imag)
                assert abs(oracle_lib.ylm_xyz(l, m, *v) - ref) < 1e-12, (l, m)


# --- a13: cubic spline -------------------------------------------------------------------

def test_spline_reproduces_cubics_and_matches_scipy(oracle_lib):
    rng = np.random.default_rng(7)
    x = np.sort(rng.uniform(0, 5, 12))
    y = 1.5 - 0.7 * x + 0.3 * x ** 2 - 0.11 * x ** 3
    c = oracle_lib.spline(x, y)
    xs = np.linspace(x[0], x[-1] - 1e-9, 57)
    val = np.array([oracle_lib.spline_eval(x, y, c, t) for t in xs])
    assert np.abs(val - (1.5 - 0.7 * xs + 0.3 * xs ** 2 - 0.11 * xs ** 3)).max() < 1e-10
    y2 = np.sin(2 * x)
    c2 = oracle_lib.spline(x, y2)
    cs = CubicSpline(x, y2, bc_type="not-a-knot")
    val2 = np.array([oracle_lib.spline_eval(x, y2, c2, t) for t in xs])
    assert np.abs(val2 - cs(xs)).max() < 1e-10
    # clamped ends (src/cspline.cpp:156-161)
    assert oracle_lib.spline_eval(x, y2, c2, x[0] - 1.0) == y2[0]
    assert oracle_lib.spline_eval(x, y2, c2, x[-1]) == y2[-1]
    assert oracle_lib.spline_eval(x, y2, c2, x[-1] + 3.0) == y2[-1]


def test_spline_rejects_bad_input(oracle_lib):
    with pytest.raises(ValueError):
        oracle_lib.spline(np.array([0.0, 1.0]), np.array([1.0, 2.0]))
    with pytest.raises(ValueError):
        oracle_lib.spline(np.array([0.0, 2.0, 1.0]), np.array([1.0, 2.0, 3.0]))


# --- a12: radial Poisson ------------------------------------------------------------------

def _multipole_U(l, a, r):
    """U_l(r) = r V_l(r) for the density r^l exp(-a r^2), closed form via incomplete gammas."""
    def lower(n):  # int_0^r s^n e^{-a s^2} ds
        return 0.5 * a ** (-(n + 1) / 2) * special.gamma((n + 1) / 2) * special.gammainc((n + 1) / 2, a * r * r)

    def upper(n):
        return 0.5 * a ** (-(n + 1) / 2) * special.gamma((n + 1) / 2) * special.gammaincc((n + 1) / 2, a * r * r)

    V = 4 * np.pi / (2 * l + 1) * (r ** (-l - 1) * lower(2 * l + 2) + r ** l * upper(1))
    return r * V


@pytest.mark.parametrize("l", [0, 1, 2])
def test_poisson_gaussian_multipoles(oracle_lib, l):
    a = 1.1
    errs = []
    for N in (30, 60, 90):
        x = np.cos(np.arange(1, N + 1) * np.pi / (N + 1))
        r = (1 + x) / (1 - x)
        M_ = oracle_lib.poisson_matrix(r, l)
        rho = r ** l * np.exp(-a * r * r)
        g = np.zeros(N + 2)
        if l == 0:  # boundary U(r -> inf) = sqrt(4 pi) q with q the charge
            q = integrate.quad(lambda s: 4 * np.pi * s * s * np.exp(-a * s * s), 0, np.inf)[0]
            g[0] = q  # rho_00 = rho here, so U_00 = r V with V -> q / r
        g[1:N + 1] = -4 * np.pi * r * rho
        U = oracle_lib.lu_solve(M_, g)[1:N + 1]
        ref = _multipole_U(l, a, r)
        inner = r < 8
        errs.append(np.abs(U[inner] - ref[inner]).max() / np.abs(ref[inner]).max())
    # the FD discretisation error falls steeply with N
    assert errs[-1] < 5e-6 and errs[0] > 10 * errs[-1]


def test_lu_solve_against_numpy(oracle_lib):
    rng = np.random.default_rng(11)
    A = rng.standard_normal((9, 9))
    b = rng.standard_normal(9)
    assert np.abs(oracle_lib.lu_solve(A, b) - np.linalg.solve(A, b)).max() < 1e-12


# --- a14/a15: Hartree J against analytic ERIs ---------------------------------------------

def test_grid_J_converges_to_eri_J(oracle_lib):
    from oracle import oracle as O
    mol = M.h2o("sto3g")
    S, T, V = O.one_electron(mol)
    P = core_guess(mol, S, T, V)
    J_eri = O.eri_J(mol, P)
    errs = []
    for nrad, nang, lmax in ((30, 50, 4), (60, 146, 8)):
        g = O.Grid(mol, nrad, nang, lmax)
        g.set_density(P)
        g.correct_densities()
        errs.append(np.abs(g.hartree() - J_eri).max())
    # J = (1/2) Phi^T diag(w V) Phi with V the potential of rho = 2 phi^T P phi equals
    # sum_kl P_kl (ij|kl) (the reference's F = H + 2J convention); |J| ~ 9.4 here
    assert errs[-1] < 1e-3 and errs[0] > 10 * errs[-1]


def test_eri_known_value(oracle_lib):
    # (ss|ss) of four normalised unit-exponent s primitives at one centre is 2/sqrt(pi);
    # the reference's prefactor and norms both use pi = 3.14159265359
    from oracle import oracle as O
    mol = M.Molecule([1], [[0, 0, 0]], {1: [(0, [1.0], [1.0])]})
    J = O.eri_J(mol, np.array([[1.0]]))
    # the Boys argument is clamped to >= 1e-8 (src/gamma.cpp:42-43): F_0(1e-8) = 1 - 3.3e-9
    x = 1e-8
    f0 = math.sqrt(math.pi) * math.erf(math.sqrt(x)) / (2 * math.sqrt(x))
    assert abs(J[0, 0] - 2 / math.sqrt(3.14159265359) * f0) < 1e-13
    assert abs(oracle_lib.boys(0, 0.0) - f0) < 1e-14


def test_boys_function(oracle_lib):
    for m in range(5):
        for x in (1e-3, 0.5, 3.0, 17.0, 45.0):
            ref = integrate.quad(lambda t: t ** (2 * m) * np.exp(-x * t * t), 0, 1, epsabs=0, epsrel=1e-13)[0]
            assert abs(oracle_lib.boys(m, x) - ref) < 1e-12 * max(ref, 1e-3)


# --- a9: XC matrix ------------------------------------------------------------------------

def test_xc_symmetry_and_gradient(oracle_lib):
    from oracle import oracle as O
    mol = M.h2o("sto3g")
    g = O.Grid(mol, 30, 50, 2)
    P = random_density(g.nbf, 4)
    g.set_density(P)  # no renormalisation: E_xc is a function of P alone
    XC, e0 = g.xc()
    assert np.abs(XC - XC.T).max() < 1e-13 * np.abs(XC).max()
    h = 1e-5
    for i, j in ((0, 0), (1, 3), (2, 6)):
        D = np.zeros_like(P)
        D[i, j] += h
        D[j, i] += h if i != j else 0.0
        g.set_density(P + D)
        ep = g.xc()[1]
        g.set_density(P - D)
        em = g.xc()[1]
        fd = (ep - em) / (2 * h)
        analytic = 2 * XC[i, j] * (2 if i != j else 1)
        assert abs(fd - analytic) < 1e-6 * max(1.0, abs(analytic))


def test_total_energy_translation_invariant(oracle_lib):
    from oracle import oracle as O
    mol = M.h2o("sto3g")
    shifted = M.Molecule(mol.Z, mol.xyz + np.array([0.37, -1.1, 2.3]), mol.basis)
    e = []
    for m in (mol, shifted):
        g = O.Grid(m, 30, 50, 4)
        en, terms, P = g.scf(threshold=1e-6, max_iter=40)
        e.append(terms["total"])
    assert abs(e[0] - e[1]) < 1e-9


# --- a16: SCF energy assembly (src/dft.cpp:84-142, 330-366, 441-447) --------------------------
#
# E = 2 tr(PH) + 2 tr(PJ) + E_nuc + E_xc and F = H + 2J + XC.  Each term is pinned
# on its own (closed forms, grid identities, a library eigen-solver), and F
# against E through the derivative: with the renormalisation factor c held,
# dE/dP = 2 (H + 2J + c XC), which is 2F up to (c - 1) XC and the small
# asymmetry of the numerical Hartree operator.

def _oracle_h2o(nrad=50, nang=110, lmax=8):
    from oracle import oracle as O
    mol = M.h2o("sto3g")
    g = O.Grid(mol, nrad, nang, lmax)
    S, T, V = O.one_electron(mol)
    return O, mol, g, core_guess(mol, S, T, V)


def test_scf_energy_gradient_is_twice_the_fock_matrix(oracle_lib):
    O, mol, g, P0 = _oracle_h2o()
    t0, F, H, J, XC = g.evaluate(P0)
    c0 = t0["c"]
    assert abs(c0 - 1.0) < 1e-4
    rng = np.random.default_rng(1)
    checks = []
    for _ in range(3):
        D = rng.standard_normal(P0.shape)
        D = 1e-3 * (D + D.T)
        ep = g.evaluate(P0 + D, hold_c=c0)[0]["total"]
        em = g.evaluate(P0 - D, hold_c=c0)[0]["total"]
        fd = (ep - em) / 2
        checks.append((fd, D))
        assert abs(fd - np.sum(D * 2 * F)) < 2e-3 * abs(fd), (fd, np.sum(D * 2 * F))
    # a dropped factor on J or XC, or a doubled H, in either E or F would fail the
    # comparison above by far more than its tolerance
    for wrong in (H + J + XC, H + 2 * J + 2 * XC, 2 * H + 2 * J + XC, H + 2 * J):
        assert max(abs(fd - np.sum(D * 2 * wrong)) / abs(fd) for fd, D in checks) > 0.05


def test_scf_hartree_energy_equals_pointwise_integral(oracle_lib):
    # 2 tr(PJ) with J = (1/2) Phi^T diag(w V) Phi is sum_p w_p V_p rho_p / (2c): the
    # same number through the grid values instead of the matrix
    O, mol, g, P0 = _oracle_h2o(30, 50, 4)
    t, F, H, J, XC = g.evaluate(P0)
    _, w, _ = g.points()
    V = g.hartree_parts()["V"]
    rho = g.densities()
    assert abs(t["J"] - np.sum(w * V * rho) / (2 * t["c"])) < 1e-11 * t["J"]
    assert abs(t["J"] - 2 * np.sum(P0 * J)) < 1e-12 * t["J"]
    # E_one from the closed-form-pinned one-electron matrices, E_xc from the grid
    S, T, Vn = O.one_electron(mol)
    assert abs(t["one"] - 2 * np.sum(P0 * (T + Vn))) < 1e-12 * abs(t["one"])
    exd, _ = g.xc_parts()
    assert abs(t["xc"] - np.sum(w * exd)) < 1e-12 * abs(t["xc"])
    assert abs(t["total"] - (t["one"] + t["J"] + t["xc"] + t["nuc"])) < 1e-12 * abs(t["total"])


def test_scf_nuclear_repulsion(oracle_lib):
    from oracle import oracle as O
    R = 1.4
    h2 = M.Molecule([1, 1], [[0, 0, 0], [0, 0, R]], "sto3g")
    g = O.Grid(h2, 10, 26, 2)
    _, terms, _ = g.scf(threshold=1.0, min_iter=0, max_iter=1)
    assert abs(terms["nuc"] - 1.0 / R) < 1e-15
    mol = M.h2o("sto3g")
    g = O.Grid(mol, 10, 26, 2)
    _, terms, _ = g.scf(threshold=1.0, min_iter=0, max_iter=1)
    Z, x = np.array(mol.Z, float), mol.xyz
    e = sum(Z[i] * Z[j] / np.linalg.norm(x[i] - x[j]) for i in range(3) for j in range(i + 1, 3))
    assert abs(terms["nuc"] - e) < 1e-13 * e


def test_scf_fixed_point_diagonalises_fock(oracle_lib):
    # at convergence the mixed P is a fixed point of its own aufbau step: idempotent
    # in the S metric, nelec/2 electrons, and commuting with F (F P S = S P F)
    O, mol, g, _ = _oracle_h2o(30, 50, 4)
    en, terms, P = g.scf(threshold=1e-11, min_iter=3, max_iter=200)
    assert abs(en[-1] - en[-2]) <= 1e-11
    t, F, H, J, XC = g.evaluate(P)
    assert abs(t["total"] - terms["total"]) < 1e-9
    S, _, _ = O.one_electron(mol)
    assert np.abs(P @ S @ P - P).max() < 1e-5
    assert abs(np.trace(P @ S) - mol.nelec / 2) < 1e-5
    comm = F @ P @ S - S @ P @ F
    assert np.abs(comm).max() < 1e-5 * np.abs(F).max()


def test_oracle_eigh_matches_numpy(oracle_lib):
    from oracle import oracle as O
    rng = np.random.default_rng(5)
    for n in (1, 3, 7, 20):
        A = rng.standard_normal((n, n))
        A = A + A.T
        w, V = O.eigh(A)
        assert np.abs(w - np.linalg.eigvalsh(A)).max() < 1e-12 * max(1.0, np.abs(w).max())
        assert np.all(np.diff(w) >= 0)
        assert np.abs(V.T @ V - np.eye(n)).max() < 1e-12
        assert np.abs(A @ V - V * w).max() < 1e-11 * max(1.0, np.abs(w).max())
    # degenerate spectrum: the eigenvalues, and the projector on each cluster
    Q, _ = np.linalg.qr(rng.standard_normal((7, 7)))
    lam = np.array([-1.0, -1.0, 0.5, 2.0, 2.0, 2.0, 3.0])
    A = Q @ np.diag(lam) @ Q.T
    w, V = O.eigh(A)
    assert np.abs(w - lam).max() < 1e-13
    for sel in ([0, 1], [2], [3, 4, 5], [6]):
        Pm = V[:, sel] @ V[:, sel].T
        Qm = Q[:, sel] @ Q[:, sel].T
        assert np.abs(Pm - Qm).max() < 1e-12


def _prim(Z, xyz, L, alpha):
    return M.Molecule(Z, xyz, {z: [(L, [alpha], [1.0])] for z in set(Z)})


@pytest.mark.parametrize("alpha", [0.3, 1.0, 4.7])
def test_one_electron_closed_forms(oracle_lib, alpha):
    from oracle import oracle as O
    Z = 7
    # one normalised s primitive on its own nucleus: T = 3a/2, V = -2Z sqrt(2a/pi) F_0(0),
    # where the Boys argument 0 is clamped to 1e-8 (src/gamma.cpp:42-43)
    f0 = math.sqrt(math.pi) * math.erf(1e-4) / (2 * 1e-4)
    S, T, V = O.one_electron(_prim([Z], [[0.1, -0.2, 0.3]], 0, alpha))
    assert abs(S[0, 0] - 1.0) < 1e-11
    assert abs(T[0, 0] - 1.5 * alpha) < 1e-11 * alpha
    v_ss = -2 * Z * math.sqrt(2 * alpha / math.pi)
    assert abs(V[0, 0] - v_ss * f0) < 1e-11 * abs(v_ss)
    # p primitives: T = 5a/2, V = -(4/3) Z sqrt(2a/pi) (to the clamp's 1e-8), the three
    # components orthogonal
    S, T, V = O.one_electron(_prim([Z], [[0.0, 0.0, 0.0]], 1, alpha))
    assert np.abs(S - np.eye(3)).max() < 1e-11
    assert np.abs(T - 2.5 * alpha * np.eye(3)).max() < 1e-11 * alpha
    v_pp = -4 / 3 * Z * math.sqrt(2 * alpha / math.pi)
    assert np.abs(V - v_pp * np.eye(3)).max() < 1e-8 * abs(v_pp)
    # two s primitives on two nuclei: overlap and nuclear attraction in closed form
    # (Gaussian product theorem, F_0(t) = sqrt(pi/t) erf(sqrt t) / 2)
    b, R = 0.8, 1.3
    mol = M.Molecule([1, 3], [[0, 0, 0], [0, 0, R]], {1: [(0, [alpha], [1.0])], 3: [(0, [b], [1.0])]})
    S, T, V = O.one_electron(mol)
    g = alpha + b
    K = math.exp(-alpha * b * R * R / g)
    assert abs(S[0, 1] - (2 * math.sqrt(alpha * b) / g) ** 1.5 * K) < 1e-12
    Na, Nb = (2 * alpha / math.pi) ** 0.75, (2 * b / math.pi) ** 0.75
    Pz = b * R / g
    v = 0.0
    for Zc, zc in ((1, 0.0), (3, R)):
        t = g * (Pz - zc) ** 2
        f0 = 1.0 if t < 1e-15 else 0.5 * math.sqrt(math.pi / t) * math.erf(math.sqrt(t))
        v -= Zc * 2 * math.pi / g * K * Na * Nb * f0
    assert abs(V[0, 1] - v) < 1e-10 * abs(v)
    # kinetic: <a|-1/2 lap|b> = ab/g (3 - 2 ab R^2/g) S_ab
    assert abs(T[0, 1] - alpha * b / g * (3 - 2 * alpha * b * R * R / g) * S[0, 1]) < 1e-12
