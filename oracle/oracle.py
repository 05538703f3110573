"""ctypes binding of the oracle (build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's
CPU-baseline leg, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def build():
    subprocess.run(["make", "-C", _HERE], check=True, capture_output=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_grid_create.restype = C.c_void_p
        L.or_grid_create.argtypes = [C.c_int, _I, _D, C.c_int, _I, _I, _I, _I, _D, _D,
                                     C.c_int, C.c_int, _D, C.c_int]
        L.or_grid_create_cells.restype = C.c_void_p
        L.or_grid_create_cells.argtypes = [C.c_int, _I, _D, C.c_int, _I, _I, _I, _I, _D, _D,
                                           C.c_int, C.c_int, _D, C.c_int, C.c_int]
        L.or_grid_destroy.argtypes = [C.c_void_p]
        L.or_grid_nbf.argtypes = [C.c_void_p]
        L.or_grid_npts.argtypes = [C.c_void_p]
        L.or_grid_npts.restype = C.c_long
        for name in ("or_grid_points",):
            getattr(L, name).argtypes = [C.c_void_p, _D, _D, _D]
        for name in ("or_grid_amplitudes", "or_grid_radial", "or_grid_set_density",
                     "or_grid_densities", "or_grid_cell_densities", "or_grid_hartree"):
            getattr(L, name).argtypes = [C.c_void_p, _D]
        L.or_grid_correct_densities.argtypes = [C.c_void_p]
        L.or_grid_calculate_density.argtypes = [C.c_void_p]
        L.or_grid_calculate_density.restype = C.c_double
        L.or_grid_hartree_parts.argtypes = [C.c_void_p, _D, _D, _D, _D, _D]
        L.or_grid_hartree_times.argtypes = [C.c_void_p, _D]
        L.or_grid_xc.argtypes = [C.c_void_p, _D, _D]
        L.or_grid_xc_parts.argtypes = [C.c_void_p, _D, _D]
        L.or_gto_norm.argtypes = [C.c_double, C.c_int, C.c_int, C.c_int]
        L.or_gto_norm.restype = C.c_double
        L.or_sh_prefactor.argtypes = [C.c_int, C.c_int]
        L.or_sh_prefactor.restype = C.c_double
        L.or_spherical_harmonic.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
        L.or_spherical_harmonic.restype = C.c_double
        L.or_ylm_xyz.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
        L.or_ylm_xyz.restype = C.c_double
        L.or_becke_cutoff.argtypes = [C.c_double]
        L.or_becke_cutoff.restype = C.c_double
        L.or_xalpha.argtypes = [C.c_long, _D, _D, _D, _D, _D]
        L.or_vwn.argtypes = [C.c_long, _D, _D, _D, _D, _D]
        L.or_spline_build.argtypes = [C.c_int, _D, _D, _D]
        L.or_spline_eval.argtypes = [C.c_int, _D, _D, _D, C.c_double]
        L.or_spline_eval.restype = C.c_double
        L.or_poisson_matrix.argtypes = [C.c_int, C.c_int, _D, _D]
        L.or_lu_solve.argtypes = [C.c_int, _D, _D, _D]
        L.or_one_electron.argtypes = [C.c_int, _I, _D, C.c_int, _I, _I, _I, _I, _D, _D, _D, _D, _D]
        L.or_boys.argtypes = [C.c_double, C.c_double]
        L.or_boys.restype = C.c_double
        L.or_eri_J.argtypes = [C.c_int, _D, C.c_int, _I, _I, _I, _I, _D, _D, _D, _D]
        L.or_eigh.argtypes = [C.c_int, _D, _D, _D]
        L.or_scf.argtypes = [C.c_void_p, C.c_int, _I, _D, C.c_int, _I, _I, _I, _I, _D, _D,
                             C.c_double, C.c_int, C.c_int, C.c_double, _D, _D, _D]
        L.or_scf.restype = C.c_int
        L.or_scf_eri.argtypes = L.or_scf.argtypes
        L.or_scf_eri.restype = C.c_int
        L.or_scf_evaluate.argtypes = [C.c_void_p, C.c_int, _I, _D, C.c_int, _I, _I, _I, _I, _D, _D,
                                      _D, C.c_double, _D, _D, _D, _D, _D]
        L.or_scf_evaluate.restype = C.c_int
        L.or_rectangular_points.argtypes = [C.c_double, C.c_uint, _D]
        L.or_points_density.argtypes = [C.c_int, _D, C.c_int, _I, _I, _I, _I, _D, _D, C.c_long, _D, _D,
                                        _D, _D, _D]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_D)


def _i(a):
    return a.ctypes.data_as(_I)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Grid:
    """Oracle molecular grid for a dftcxx_amd.molecule.Molecule."""

    def __init__(self, mol, nrad, nang, lmax=8, leb=None, ncell=None):
        """ncell: build points in only the first ncell cells (a CPU timing sample)."""
        from dftcxx_amd import lebedev
        self.mol = mol
        self.sh = mol.shells()
        self.nrad, self.nang, self.lmax = nrad, nang, lmax
        self.nlm = (lmax + 1) ** 2
        self._leb = _f64(lebedev.rule(nang) if leb is None else leb)
        s = self.sh
        self._keep = [_i32(mol.Z), _f64(mol.xyz), _i32(s["atom"]), _i32(s["L"]), _i32(s["nprim"]),
                      _i32(s["offset"]), _f64(s["alpha"]), _f64(s["coef"])]
        Z, xyz, at, L, npr, off, al, co = self._keep
        self.ncell = mol.natoms if ncell is None else int(ncell)
        self.h = lib().or_grid_create_cells(mol.natoms, _i(Z), _d(xyz), len(at), _i(at), _i(L), _i(npr),
                                            _i(off), _d(al), _d(co), nrad, nang, _d(self._leb), lmax,
                                            self.ncell)
        self.nbf = lib().or_grid_nbf(self.h)
        self.npts = lib().or_grid_npts(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().or_grid_destroy(self.h)
            except Exception:  # interpreter shutdown: the library is already gone
                pass
            self.h = None

    def points(self):
        xyz = np.empty((self.npts, 3))
        w = np.empty(self.npts)
        wb = np.empty(self.npts)
        lib().or_grid_points(self.h, _d(xyz), _d(w), _d(wb))
        return xyz, w, wb

    def amplitudes(self):
        phi = np.empty((self.npts, self.nbf))
        lib().or_grid_amplitudes(self.h, _d(phi))
        return phi

    def radial(self):
        r = np.empty(self.nrad)
        lib().or_grid_radial(self.h, _d(r))
        return r

    def set_density(self, P):
        P = _f64(P)
        lib().or_grid_set_density(self.h, _d(P))

    def correct_densities(self):
        lib().or_grid_correct_densities(self.h)

    def calculate_density(self):
        return lib().or_grid_calculate_density(self.h)

    def densities(self):
        r = np.empty(self.npts)
        lib().or_grid_densities(self.h, _d(r))
        return r

    def cell_densities(self):
        q = np.empty(self.mol.natoms)
        lib().or_grid_cell_densities(self.h, _d(q))
        return q

    def hartree(self):
        J = np.empty((self.nbf, self.nbf))
        lib().or_grid_hartree(self.h, _d(J))
        return J

    def hartree_times(self):
        """Seconds of the last hartree(): projection, Poisson+own cell+splines, cross-atom V, J."""
        t = np.zeros(4)
        lib().or_grid_hartree_times(self.h, _d(t))
        return t

    def hartree_parts(self):
        na = self.mol.natoms
        rho_lm = np.empty((na, self.nrad, self.nlm))
        U = np.empty_like(rho_lm)
        vcell = np.empty(self.npts)
        v = np.empty(self.npts)
        spl = np.empty((na, self.nlm, self.nrad - 1, 4))
        lib().or_grid_hartree_parts(self.h, _d(rho_lm), _d(U), _d(vcell), _d(v), _d(spl))
        return {"rho_lm": rho_lm, "U_lm": U, "vcell": vcell, "V": v, "spline": spl}

    def xc(self):
        XC = np.empty((self.nbf, self.nbf))
        e = np.zeros(1)
        lib().or_grid_xc(self.h, _d(XC), _d(e))
        return XC, float(e[0])

    def xc_parts(self):
        exc = np.empty(self.npts)
        wva = np.empty(self.npts)
        lib().or_grid_xc_parts(self.h, _d(exc), _d(wva))
        return exc, wva

    def scf(self, threshold=1e-6, min_iter=3, max_iter=100, mixing=0.5, eri=False):
        """Run the reference SCF on this grid (J from the grid, or from the
        two-electron integrals with eri=True); returns (energies, terms, P)."""
        Z, xyz, at, L, npr, off, al, co = self._keep
        en = np.zeros(max_iter)
        terms = np.zeros(5)
        P = np.zeros((self.nbf, self.nbf))
        fn = lib().or_scf_eri if eri else lib().or_scf
        it = fn(self.h, self.mol.natoms, _i(Z), _d(xyz), len(at), _i(at), _i(L), _i(npr),
                _i(off), _d(al), _d(co), threshold, min_iter, max_iter, mixing,
                _d(en), _d(terms), _d(P))
        return en[:it], dict(zip(("total", "one", "J", "xc", "nuc"), terms)), P

    def evaluate(self, P, hold_c=0.0):
        """The SCF's grid work and energy at P (the code or_scf runs per iteration).

        hold_c > 0 scales rho by that factor instead of renormalising it.
        Returns (terms dict with total, one, J, xc, nuc, c; F; H; J; XC)."""
        Z, xyz, at, L, npr, off, al, co = self._keep
        n = self.nbf
        Pc = _f64(np.asarray(P).T)
        terms = np.zeros(6)
        F, H, J, XC = (np.empty((n, n)) for _ in range(4))
        lib().or_scf_evaluate(self.h, self.mol.natoms, _i(Z), _d(xyz), len(at), _i(at), _i(L), _i(npr),
                              _i(off), _d(al), _d(co), _d(Pc), float(hold_c), _d(terms), _d(F), _d(H),
                              _d(J), _d(XC))
        t = dict(zip(("total", "one", "J", "xc", "nuc", "c"), terms))
        return t, F.T.copy(), H.T.copy(), J.T.copy(), XC.T.copy()


def one_electron(mol):
    s = mol.shells()
    n = s["nbf"]
    S, T, V = (np.empty((n, n)) for _ in range(3))
    keep = [_i32(mol.Z), _f64(mol.xyz), _i32(s["atom"]), _i32(s["L"]), _i32(s["nprim"]),
            _i32(s["offset"]), _f64(s["alpha"]), _f64(s["coef"])]
    Z, xyz, at, L, npr, off, al, co = keep
    lib().or_one_electron(mol.natoms, _i(Z), _d(xyz), len(at), _i(at), _i(L), _i(npr), _i(off),
                          _d(al), _d(co), _d(S), _d(T), _d(V))
    return S, T, V


def eri_J(mol, P):
    s = mol.shells()
    n = s["nbf"]
    J = np.empty((n, n))
    keep = [_f64(mol.xyz), _i32(s["atom"]), _i32(s["L"]), _i32(s["nprim"]), _i32(s["offset"]),
            _f64(s["alpha"]), _f64(s["coef"]), _f64(P)]
    xyz, at, L, npr, off, al, co, Pc = keep
    lib().or_eri_J(mol.natoms, _d(xyz), len(at), _i(at), _i(L), _i(npr), _i(off), _d(al), _d(co),
                   _d(Pc), _d(J))
    return J


def eigh(A):
    """Ascending eigenvalues and eigenvectors (columns) of a symmetric matrix."""
    A = _f64(np.asarray(A).T)
    n = A.shape[0]
    w = np.empty(n)
    V = np.empty((n, n))  # filled column-major
    lib().or_eigh(n, _d(A), _d(w), _d(V))
    return w, V.T.copy()


def lda(rho):
    """Reference LDA at rho_a = rho_b = rho/2: (ex+ec, vxa, vca)."""
    rho = _f64(rho)
    n = len(rho)
    a = rho * 0.5
    b = rho * 0.5
    ex, vxa, vxb, ec, vca, vcb = (np.empty(n) for _ in range(6))
    lib().or_xalpha(n, _d(a), _d(b), _d(ex), _d(vxa), _d(vxb))
    lib().or_vwn(n, _d(a), _d(b), _d(ec), _d(vca), _d(vcb))
    return ex, ec, vxa, vca


def spline(x, y):
    x, y = _f64(x), _f64(y)
    c = np.empty((len(x) - 1, 4))
    if lib().or_spline_build(len(x), _d(x), _d(y), _d(c)) != 0:
        raise ValueError("invalid spline input")
    return c


def spline_eval(x, y, c, xx):
    x, y, c = _f64(x), _f64(y), _f64(c)
    return lib().or_spline_eval(len(x), _d(x), _d(y), _d(c), float(xx))


def poisson_matrix(r, l):
    r = _f64(r)
    N = len(r)
    M = np.empty((N + 2) * (N + 2))
    lib().or_poisson_matrix(N, l, _d(r), _d(M))
    return M.reshape(N + 2, N + 2).T  # column-major -> row-major view of M


def lu_solve(A, b):
    A = _f64(np.asarray(A).T)  # column-major
    b = _f64(b)
    x = np.empty(len(b))
    lib().or_lu_solve(len(b), _d(A), _d(b), _d(x))
    return x


def ylm_xyz(l, m, x, y, z):
    return lib().or_ylm_xyz(l, m, x, y, z)


def becke_cutoff(mu):
    return lib().or_becke_cutoff(mu)


def gto_norm(alpha, l, m, n):
    return lib().or_gto_norm(alpha, l, m, n)


def boys(m, x):
    return lib().or_boys(m, x)


def rectangular_points(size, dp):
    """RectangularGrid::build_grid positions, (dp^3, 3)."""
    out = np.empty(3 * dp ** 3)
    lib().or_rectangular_points(float(size), int(dp), _d(out))
    return out.reshape(-1, 3)


def points_density(mol, pts, P):
    """(rho [n], grad rho [n, 3], basis gradients [n, 3, nbf]) at the points, the
    reference's GridPoint::set_density / set_gradient with P symmetric."""
    s = mol.shells()
    nbf = s["nbf"]
    pts = _f64(np.asarray(pts).reshape(-1, 3))
    n = len(pts)
    Pc = _f64(np.asarray(P).T)
    rho, grad, dphi = np.empty(n), np.empty((n, 3)), np.empty((n, 3, nbf))
    keep = [_f64(mol.xyz), _i32(s["atom"]), _i32(s["L"]), _i32(s["nprim"]), _i32(s["offset"]),
            _f64(s["alpha"]), _f64(s["coef"])]
    xyz, at, L, npr, off, al, co = keep
    lib().or_points_density(mol.natoms, _d(xyz), len(at), _i(at), _i(L), _i(npr), _i(off), _d(al), _d(co), n,
                            _d(pts), _d(Pc), _d(rho), _d(grad), _d(dphi))
    return rho, grad, dphi
