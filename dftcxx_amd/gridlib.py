"""Python binding of the dftgrid C ABI (include/dftgrid.h).

:class:`MolecularGrid` mirrors the reference's ``MolecularGrid`` class
(src/moleculargrid.h:48-190): same method names, same meaning of arguments,
errors raised as exceptions.  Everything runs in libdftgrid.so on the GPU; if
the library is missing or fails to load, :func:`lib` raises -- there is no CPU
fallback.
"""
import ctypes as C
import os

import numpy as np

from . import lebedev as _leb

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libdftgrid.so")
_lib = None

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)

PREC_FP64 = 0
PREC_MIXED = 1

# every entry point declared in include/dftgrid.h
EXPORTS = (
    "dg_last_error", "dg_version", "dg_lebedev_available", "dg_create", "dg_destroy", "dg_info",
    "dg_set_stream", "dg_get_stream", "dg_set_density", "dg_correct_densities",
    "dg_calculate_density", "dg_hartree", "dg_xc", "dg_get_weights", "dg_get_becke_weights",
    "dg_get_positions", "dg_get_densities", "dg_get_amplitudes", "dg_get_hartree_potential",
    "dg_get_hartree_parts", "dg_get_xc_parts", "dg_xc_adjoint", "dg_build_device", "dg_profile_enable",
    "dg_profile_read", "dg_profile_reset", "dg_selftest_mfma", "dg_mfma_probe", "dg_mfma_probe_f32", "dg_scf_run", "dg_one_electron",
    "dg_eigh", "dg_batch_create", "dg_batch_destroy", "dg_batch_info", "dg_batch_layout",
    "dg_batch_handle", "dg_batch_set_stream", "dg_batch_build_device", "dg_batch_build",
    "dg_batch_autotune", "dg_batch_scf_run", "dg_input_read", "dg_input_destroy", "dg_input_molecule",
    "dg_input_basis", "dg_input_grid", "dg_input_value", "dg_get_amplitude_gradients",
    "dg_get_density_gradients", "dg_density_at_points", "dg_density_at_points_noalloc", "dg_build_fock_device",
    "dg_build_xc_device", "dg_eri_create", "dg_eri_destroy", "dg_eri_count", "dg_eri_get", "dg_eri_coulomb_device",
    "dg_eri_coulomb",
)


class DgError(RuntimeError):
    """A dftgrid call returned an error code."""

    def __init__(self, code, msg):
        super().__init__(f"dftgrid error {code}: {msg}")
        self.code = code


class _Molecule(C.Structure):
    _fields_ = [("natoms", C.c_int), ("Z", _I), ("xyz", _D)]


class _Basis(C.Structure):
    _fields_ = [("nshell", C.c_int), ("atom", _I), ("L", _I), ("nprim", _I), ("offset", _I),
                ("alpha", _D), ("coef", _D)]


class _GridParams(C.Structure):
    _fields_ = [("n_radial", C.c_int), ("n_angular", C.c_int), ("lmax", C.c_int),
                ("precision", C.c_int), ("screen", C.c_double)]


class _ScfParams(C.Structure):
    _fields_ = [("threshold", C.c_double), ("min_iter", C.c_int), ("max_iter", C.c_int),
                ("mixing", C.c_double), ("density_step", C.c_int), ("hartree", C.c_int)]


class _ScfResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("energy", C.c_double), ("e_one", C.c_double),
                ("e_j", C.c_double), ("e_xc", C.c_double), ("e_nuc", C.c_double),
                ("nelec", C.c_double), ("seconds_setup", C.c_double),
                ("seconds_iter", C.c_double), ("converged", C.c_int), ("delta", C.c_double),
                ("device_step", C.c_int), ("host_fallbacks", C.c_int)]


def lib():
    """Load libdftgrid.so (raises if it is missing: no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing; build it with `python -m dftcxx_amd.build`")
        # PyTorch ships its own libamdhip64.so.7 with the same soname as ROCm's; the
        # first one loaded serves the whole process.  Load torch's first when torch
        # is installed, so that torch can still initialise its device afterwards.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        L.dg_last_error.restype = C.c_char_p
        L.dg_create.argtypes = [C.POINTER(_Molecule), C.POINTER(_Basis), C.POINTER(_GridParams),
                                C.POINTER(C.c_void_p)]
        L.dg_destroy.argtypes = [C.c_void_p]
        L.dg_info.argtypes = [C.c_void_p, _I, C.POINTER(C.c_long), _I]
        L.dg_set_stream.argtypes = [C.c_void_p, C.c_void_p]
        L.dg_get_stream.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.dg_set_density.argtypes = [C.c_void_p, _D, C.c_int]
        L.dg_correct_densities.argtypes = [C.c_void_p, _D]
        L.dg_calculate_density.argtypes = [C.c_void_p, _D]
        L.dg_hartree.argtypes = [C.c_void_p, _D]
        L.dg_xc.argtypes = [C.c_void_p, _D, _D]
        for n in ("dg_get_weights", "dg_get_becke_weights", "dg_get_positions", "dg_get_densities",
                  "dg_get_amplitudes", "dg_get_hartree_potential", "dg_get_amplitude_gradients",
                  "dg_get_density_gradients"):
            getattr(L, n).argtypes = [C.c_void_p, _D]
        L.dg_get_hartree_parts.argtypes = [C.c_void_p, _D, _D, _D]
        L.dg_get_xc_parts.argtypes = [C.c_void_p, _D, _D]
        L.dg_xc_adjoint.argtypes = [C.c_void_p, C.c_int, _D, _D]
        L.dg_build_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.dg_build_fock_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.dg_build_xc_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.dg_eri_create.argtypes = [C.POINTER(_Molecule), C.POINTER(_Basis), C.POINTER(C.c_void_p)]
        L.dg_eri_destroy.argtypes = [C.c_void_p]
        L.dg_eri_count.argtypes = [C.c_void_p, C.POINTER(C.c_long), _I]
        L.dg_eri_get.argtypes = [C.c_void_p, _D]
        L.dg_eri_coulomb_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.dg_eri_coulomb.argtypes = [C.c_void_p, _D, _D]
        L.dg_profile_enable.argtypes = [C.c_void_p, C.c_int]
        L.dg_profile_read.argtypes = [C.c_void_p, C.c_int, C.c_char_p, _D, _I, _I]
        L.dg_profile_reset.argtypes = [C.c_void_p]
        L.dg_selftest_mfma.argtypes = [_I]
        L.dg_mfma_probe.argtypes = [_D, _D, _D]
        L.dg_mfma_probe_f32.argtypes = [C.POINTER(C.c_float)] * 3
        L.dg_scf_run.argtypes = [C.POINTER(_Molecule), C.POINTER(_Basis), C.POINTER(_GridParams),
                                 C.POINTER(_ScfParams), C.POINTER(_ScfResult), _D, _D]
        L.dg_one_electron.argtypes = [C.POINTER(_Molecule), C.POINTER(_Basis), _D, _D, _D]
        L.dg_eigh.argtypes = [C.c_int, _D, _D, _D]
        L.dg_batch_create.argtypes = [C.c_int, C.POINTER(_Molecule), C.POINTER(_Basis),
                                      C.POINTER(_GridParams), C.POINTER(C.c_void_p)]
        L.dg_batch_destroy.argtypes = [C.c_void_p]
        L.dg_batch_info.argtypes = [C.c_void_p, _I, C.POINTER(C.c_long), _I, C.POINTER(C.c_long)]
        L.dg_batch_layout.argtypes = [C.c_void_p, _I, C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.dg_batch_handle.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.dg_batch_set_stream.argtypes = [C.c_void_p, C.c_void_p]
        L.dg_batch_build_device.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        L.dg_batch_build.argtypes = [C.c_void_p, _D, _D, _D, _D]
        L.dg_batch_autotune.argtypes = [C.c_void_p, _I, _I, _D]
        L.dg_batch_scf_run.argtypes = [C.c_int, C.POINTER(_Molecule), C.POINTER(_Basis),
                                       C.POINTER(_GridParams), C.POINTER(_ScfParams),
                                       C.POINTER(_ScfResult), _I]
        L.dg_density_at_points.argtypes = [C.POINTER(_Molecule), C.POINTER(_Basis), C.c_long, _D, _D, _D, _D]
        L.dg_density_at_points_noalloc.argtypes = [C.POINTER(_Molecule), C.POINTER(_Basis), C.c_long, _D, _D, _D,
                                                   _D]
        L.dg_input_read.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]
        L.dg_input_destroy.argtypes = [C.c_void_p]
        L.dg_input_molecule.argtypes = [C.c_void_p, C.POINTER(_Molecule)]
        L.dg_input_basis.argtypes = [C.c_void_p, C.POINTER(_Basis)]
        L.dg_input_grid.argtypes = [C.c_void_p, C.POINTER(_GridParams), _I, _I]
        L.dg_input_value.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_char_p)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise DgError(rc, lib().dg_last_error().decode())


def _d(a):
    return a.ctypes.data_as(_D)


def _i(a):
    return a.ctypes.data_as(_I)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _c_inputs(mol, nrad, nang, lmax, precision, screen=0.0):
    s = mol.shells()
    keep = {"Z": _i32(mol.Z), "xyz": _f64(mol.xyz), "atom": _i32(s["atom"]), "L": _i32(s["L"]),
            "nprim": _i32(s["nprim"]), "offset": _i32(s["offset"]), "alpha": _f64(s["alpha"]),
            "coef": _f64(s["coef"])}
    m = _Molecule(mol.natoms, _i(keep["Z"]), _d(keep["xyz"]))
    b = _Basis(len(keep["atom"]), _i(keep["atom"]), _i(keep["L"]), _i(keep["nprim"]),
               _i(keep["offset"]), _d(keep["alpha"]), _d(keep["coef"]))
    g = _GridParams(nrad, nang, lmax, precision, screen)
    return m, b, g, keep


def selftest_mfma():
    ok = C.c_int(0)
    _check(lib().dg_selftest_mfma(C.byref(ok)))
    return bool(ok.value)


def mfma_probe(a, b):
    """Register file (4 x 64) of one FP64 16x16x4 MFMA with lane inputs a, b."""
    a, b = _f64(a), _f64(b)
    d = np.empty(256)
    _check(lib().dg_mfma_probe(_d(a), _d(b), _d(d)))
    return d.reshape(4, 64)


def mfma_probe_f32(a, b):
    """FP32 register file (4 x 64) of one FP32 16x16x4 MFMA with lane inputs a, b (64 each)."""
    F = C.POINTER(C.c_float)
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    d = np.empty(256, dtype=np.float32)
    _check(lib().dg_mfma_probe_f32(a.ctypes.data_as(F), b.ctypes.data_as(F), d.ctypes.data_as(F)))
    return d.reshape(4, 64)


class MolecularGrid:
    """GPU molecular grid with the reference's MolecularGrid call surface.

    ``MolecularGrid(mol)`` then ``set_grid_parameters(radial, lebedev_order,
    lmax)`` and ``create_grid()`` as in src/dft.cpp:57-61, where
    ``lebedev_order`` is the reference's enum index (0..10 -> 6..194 points);
    or pass ``radial_points``/``angular_points`` (a point count, e.g. 302)
    directly, which builds the grid at once.
    """

    def __init__(self, mol, radial_points=None, angular_points=None, lmax=8, precision=PREC_FP64,
                 screen=0.0):
        self.mol = mol
        self.h = None
        self.radial_points = radial_points
        self.angular_points = angular_points
        self.lmax = lmax
        self.precision = precision
        self.screen = screen
        if radial_points is not None and angular_points is not None:
            self.create_grid()

    # -- set-up (src/moleculargrid.cpp:175-261) ------------------------------
    def set_grid_parameters(self, radial_points, lebedev_order, lmax):
        self.radial_points = int(radial_points)
        self.angular_points = _leb.points_from_reference_index(int(lebedev_order))
        self.lmax = int(lmax)

    def create_grid(self):
        if self.h is not None:
            raise RuntimeError("grid already created")
        m, b, g, keep = _c_inputs(self.mol, self.radial_points, self.angular_points, self.lmax,
                                  self.precision, self.screen)
        h = C.c_void_p()
        _check(lib().dg_create(C.byref(m), C.byref(b), C.byref(g), C.byref(h)))
        self.h = h
        nbf, npts, nat = C.c_int(), C.c_long(), C.c_int()
        _check(lib().dg_info(self.h, C.byref(nbf), C.byref(npts), C.byref(nat)))
        self.nbf, self.npts, self.natoms = nbf.value, npts.value, nat.value
        self.nlm = (self.lmax + 1) ** 2

    def close(self):
        if self.h is not None:
            lib().dg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _need(self):
        if self.h is None:
            raise RuntimeError("create_grid() has not been called")

    # -- per-iteration work ---------------------------------------------------
    def set_density(self, P):
        """rho = 2 phi^T P phi at every point (src/moleculargrid.cpp:48-53)."""
        self._need()
        P = _f64(np.asarray(P).T)  # column-major
        if P.shape != (self.nbf, self.nbf):
            raise ValueError(f"P must be {self.nbf}x{self.nbf}")
        _check(lib().dg_set_density(self.h, _d(P), self.nbf))

    def correct_densities(self):
        """Scale rho to the electron count; returns the integral before (src/moleculargrid.cpp:132-146)."""
        self._need()
        before = C.c_double()
        _check(lib().dg_correct_densities(self.h, C.byref(before)))
        return before.value

    def calculate_density(self):
        self._need()
        n = C.c_double()
        _check(lib().dg_calculate_density(self.h, C.byref(n)))
        return n.value

    def calculate_hartree_potential(self):
        """J matrix from the Becke-Poisson Hartree potential (src/moleculargrid.cpp:336-389)."""
        self._need()
        J = np.empty((self.nbf, self.nbf))
        _check(lib().dg_hartree(self.h, _d(J)))
        return J.T.copy()

    def calculate_exchange_correlation(self):
        """(XC matrix, E_xc) as DFT::calculate_exchange_correlation_matrix (src/dft.cpp:394-433)."""
        self._need()
        XC = np.empty((self.nbf, self.nbf))
        e = C.c_double()
        _check(lib().dg_xc(self.h, _d(XC), C.byref(e)))
        return XC.T.copy(), e.value

    # -- getters (src/moleculargrid.cpp:61-127) ---------------------------------
    def _vec(self, fn, n):
        self._need()
        out = np.empty(n)
        _check(fn(self.h, _d(out)))
        return out

    def get_weights(self):
        return self._vec(lib().dg_get_weights, self.npts)

    def get_becke_weights(self):
        return self._vec(lib().dg_get_becke_weights, self.npts)

    def get_densities(self):
        return self._vec(lib().dg_get_densities, self.npts)

    def get_hartree_potential(self):
        return self._vec(lib().dg_get_hartree_potential, self.npts)

    def get_positions(self):
        return self._vec(lib().dg_get_positions, 3 * self.npts).reshape(self.npts, 3)

    def get_amplitudes(self):
        """Amplitudes as (npts, nbf); the reference returns the transpose (nbf x npts)."""
        return self._vec(lib().dg_get_amplitudes, self.npts * self.nbf).reshape(self.npts, self.nbf)

    def get_amplitude_gradients(self):
        """Basis-function gradients (npts, 3, nbf) (src/gridpoint.cpp:62-72)."""
        return self._vec(lib().dg_get_amplitude_gradients, self.npts * 3 * self.nbf).reshape(
            self.npts, 3, self.nbf)

    def get_density_gradients(self):
        """Gradient of the current density (npts, 3) (src/gridpoint.cpp:94-108)."""
        return self._vec(lib().dg_get_density_gradients, 3 * self.npts).reshape(self.npts, 3)

    def get_hartree_parts(self):
        self._need()
        shape = (self.natoms, self.radial_points, self.nlm)
        rho_lm, U, vcell = np.empty(shape), np.empty(shape), np.empty(self.npts)
        _check(lib().dg_get_hartree_parts(self.h, _d(rho_lm), _d(U), _d(vcell)))
        return {"rho_lm": rho_lm, "U_lm": U, "vcell": vcell}

    def get_xc_parts(self):
        self._need()
        e, wv = np.empty(self.npts), np.empty(self.npts)
        _check(lib().dg_get_xc_parts(self.h, _d(e), _d(wv)))
        return e, wv

    def xc_adjoint(self, hold_c=False, dEdP=True, grad=True):
        """(dE_xc/dP [nbf, nbf], dE_xc/dR [natoms, 3]) of the last XC evaluation.

        hold_c=True keeps the renormalisation factor c fixed (dE/dP = 2 c XC);
        False differentiates through c as well.  Either output can be skipped.
        """
        self._need()
        n = self.nbf
        dP = np.empty((n, n)) if dEdP else None
        g = np.empty((self.natoms, 3)) if grad else None
        _check(lib().dg_xc_adjoint(self.h, 1 if hold_c else 0,
                                   _d(dP) if dP is not None else None, _d(g) if g is not None else None))
        return (dP.T.copy() if dP is not None else None), g

    # -- device-pointer build and profiling ---------------------------------
    def set_stream(self, stream_ptr):
        self._need()
        _check(lib().dg_set_stream(self.h, C.c_void_p(stream_ptr)))

    def build_device(self, dP, dJ=0, dXC=0, dscalars=0):
        """Enqueue one full grid build on device pointers (ints), see dg_build_device."""
        self._need()
        _check(lib().dg_build_device(self.h, C.c_void_p(dP), C.c_void_p(dJ), C.c_void_p(dXC),
                                     C.c_void_p(dscalars)))

    def build_fock_device(self, dP, dG, dscalars=0):
        """Enqueue one build of G = 2J + XC on device pointers (ints), see dg_build_fock_device."""
        self._need()
        _check(lib().dg_build_fock_device(self.h, C.c_void_p(dP), C.c_void_p(dG), C.c_void_p(dscalars)))

    def profile_enable(self, on=True):
        self._need()
        _check(lib().dg_profile_enable(self.h, int(on)))

    def profile_reset(self):
        self._need()
        _check(lib().dg_profile_reset(self.h))

    def profile_read(self):
        self._need()
        return _profile_read(self.h)


def _profile_read(h):
    """{kernel name: (total ms, launches)} accumulated on handle h."""
    maxn = 64
    names = C.create_string_buffer(48 * maxn)
    ms = np.zeros(maxn)
    cnt = np.zeros(maxn, dtype=np.int32)
    n = C.c_int()
    _check(lib().dg_profile_read(h, maxn, names, _d(ms), _i(cnt), C.byref(n)))
    out = {}
    for i in range(min(n.value, maxn)):
        nm = names.raw[48 * i:48 * i + 48].split(b"\0", 1)[0].decode()
        out[nm] = (float(ms[i]), int(cnt[i]))
    return out


class Batch:
    """Many molecules on one grid specification, built together (dg_batch_*).

    Every SCF iteration is one hipGraph launch for the whole batch; matrices
    are passed as lists of (nbf_m, nbf_m) arrays, or packed for the device
    entry point (molecule m's column-major matrix at ``offset[m]``).
    """

    def __init__(self, mols, radial_points, angular_points, lmax=8):
        self.mols = list(mols)
        n = len(self.mols)
        ins = [_c_inputs(m, radial_points, angular_points, lmax, PREC_FP64) for m in self.mols]
        self._keep = ins
        ms = (_Molecule * n)(*[i[0] for i in ins])
        bs = (_Basis * n)(*[i[1] for i in ins])
        g = ins[0][2]
        h = C.c_void_p()
        self.b = None
        _check(lib().dg_batch_create(n, ms, bs, C.byref(g), C.byref(h)))
        self.b = h
        nm, npts, nat, packed = C.c_int(), C.c_long(), C.c_int(), C.c_long()
        _check(lib().dg_batch_info(self.b, C.byref(nm), C.byref(npts), C.byref(nat), C.byref(packed)))
        self.nmol, self.npts, self.natoms, self.packed = nm.value, npts.value, nat.value, packed.value
        self.nbf = np.empty(n, np.int32)
        self.offset = np.empty(n + 1, np.int64)
        self.point0 = np.empty(n + 1, np.int64)
        L64 = C.POINTER(C.c_long)
        _check(lib().dg_batch_layout(self.b, _i(self.nbf), self.offset.ctypes.data_as(L64),
                                     self.point0.ctypes.data_as(L64)))
        hh = C.c_void_p()
        _check(lib().dg_batch_handle(self.b, C.byref(hh)))
        self.h = hh

    def close(self):
        if self.b is not None:
            lib().dg_batch_destroy(self.b)
            self.b = None
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pack(self, mats):
        """Packed column-major vector of per-molecule matrices."""
        out = np.empty(self.packed)
        for m, A in enumerate(mats):
            out[self.offset[m]:self.offset[m + 1]] = np.asarray(A, dtype=np.float64).T.ravel()
        return out

    def unpack(self, flat):
        return [flat[self.offset[m]:self.offset[m + 1]].reshape(self.nbf[m], self.nbf[m]).T.copy()
                for m in range(self.nmol)]

    def build(self, Ps):
        """(list of J, list of XC, scalars (nmol, 4)) for the density matrices Ps."""
        P = self.pack(Ps)
        J, XC = np.empty(self.packed), np.empty(self.packed)
        sc = np.empty((self.nmol, 4))
        _check(lib().dg_batch_build(self.b, _d(P), _d(J), _d(XC), _d(sc)))
        return self.unpack(J), self.unpack(XC), sc

    def build_device(self, dP, dJ, dXC, dscalars):
        _check(lib().dg_batch_build_device(self.b, C.c_void_p(dP), C.c_void_p(dJ), C.c_void_p(dXC),
                                           C.c_void_p(dscalars)))

    def set_stream(self, stream_ptr):
        _check(lib().dg_batch_set_stream(self.b, C.c_void_p(stream_ptr)))

    def autotune(self):
        """(contraction chunk length, cross item length, their best ms) chosen on this batch."""
        c, i = C.c_int(), C.c_int()
        ms = np.zeros(2)
        _check(lib().dg_batch_autotune(self.b, C.byref(c), C.byref(i), _d(ms)))
        return c.value, i.value, ms

    def get_weights(self):
        w = np.empty(self.npts)
        _check(lib().dg_get_weights(self.h, _d(w)))
        return w

    def get_densities(self):
        r = np.empty(self.npts)
        _check(lib().dg_get_densities(self.h, _d(r)))
        return r

    def profile_enable(self, on=True):
        _check(lib().dg_profile_enable(self.h, int(on)))

    def profile_reset(self):
        _check(lib().dg_profile_reset(self.h))

    def profile_read(self):
        return _profile_read(self.h)


def one_electron(mol):
    """(S, T, V) one-electron matrices on the host (src/dft.cpp:184-201)."""
    m, b, g, keep = _c_inputs(mol, 5, 6, 0, PREC_FP64)
    n = mol.nbf()
    S, T, V = (np.empty((n, n)) for _ in range(3))
    _check(lib().dg_one_electron(C.byref(m), C.byref(b), _d(S), _d(T), _d(V)))
    return S.T.copy(), T.T.copy(), V.T.copy()


def density_at_points(mol, xyz, P, noalloc=False):
    """(rho [n], grad rho [n, 3]) at arbitrary points (bohr) on the GPU (dg_density_at_points,
    or dg_density_at_points_noalloc, which allocates no device memory)."""
    m, b, g, keep = _c_inputs(mol, 5, 6, 0, PREC_FP64)
    xyz = _f64(np.asarray(xyz).reshape(-1, 3))
    n = len(xyz)
    Pc = _f64(np.asarray(P).T)
    rho, grad = np.empty(n), np.empty((n, 3))
    fn = lib().dg_density_at_points_noalloc if noalloc else lib().dg_density_at_points
    _check(fn(C.byref(m), C.byref(b), n, _d(xyz), _d(Pc), _d(rho), _d(grad)))
    return rho, grad


class RectangularGrid:
    """The reference's RectangularGrid (src/rectangulargrid.cpp:34-95): dp^3 points
    spanning [-size/2, size/2] (x fastest), density and gradient at each on the GPU,
    and the gradient file."""

    def __init__(self, mol):
        self.mol = mol
        self.positions = np.zeros((0, 3))
        self.densities = self.gradients = None

    def build_grid(self, size, dp):
        if dp < 2:
            raise ValueError("at least two points per direction")
        t = np.arange(dp) * (size / (dp - 1)) - size / 2.0
        z, y, x = np.meshgrid(t, t, t, indexing="ij")
        self.positions = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)

    def set_density(self, P):
        self.densities, self.gradients = density_at_points(self.mol, self.positions, P)

    def write_gradient(self, filename):
        with open(filename, "w") as f:
            for p, g in zip(self.positions, self.gradients):
                f.write("%12.8f  %12.8f  %12.8f  %12.8f  %12.8f  %12.8f\n" % (p[0], p[1], p[2], g[0], g[1], g[2]))


def eigh(A):
    """Ascending eigenvalues and eigenvectors of a symmetric matrix (host)."""
    A = _f64(np.asarray(A).T)
    n = A.shape[0]
    w, Z = np.empty(n), np.empty((n, n))
    _check(lib().dg_eigh(n, _d(A), _d(w), _d(Z)))
    return w, Z.T.copy()


STEP_AUTO, STEP_HOST, STEP_DEVICE = 0, 1, 2


def scf_run(mol, nrad, nang, lmax=8, threshold=1e-6, min_iter=3, max_iter=100, mixing=0.5,
            precision=PREC_FP64, screen=0.0, density_step=STEP_AUTO, hartree=0):
    """Reference SCF (src/dft.cpp:84-142) with the grid work on the GPU; the
    density-matrix step on the host or the GPU (density_step, see dg_scf_params);
    J from the Becke-grid Poisson path or the two-electron integrals (hartree =
    HARTREE_BECKE or HARTREE_ERI).

    Returns (energies per iteration, result dict, P)."""
    m, b, g, keep = _c_inputs(mol, nrad, nang, lmax, precision, screen)
    sp = _ScfParams(threshold, min_iter, max_iter, mixing, density_step, hartree)
    res = _ScfResult()
    en = np.zeros(max(int(max_iter), 1))
    n = mol.nbf()
    P = np.zeros((n, n))
    _check(lib().dg_scf_run(C.byref(m), C.byref(b), C.byref(g), C.byref(sp), C.byref(res), _d(en), _d(P)))
    r = {k: getattr(res, k) for k, _ in _ScfResult._fields_}
    return en[:res.iterations], r, P.T.copy()


def batch_scf_run(mols, nrad, nang, lmax=8, threshold=1e-6, min_iter=3, max_iter=100, mixing=0.5,
                  density_step=STEP_AUTO):
    """The reference SCF for every molecule of a batch, one batched grid build per iteration;
    the molecules' density-matrix steps on the GPU (STEP_AUTO, STEP_DEVICE) or on host threads
    (STEP_HOST).

    Returns (list of result dicts, number of batch iterations)."""
    mols = list(mols)
    n = len(mols)
    ins = [_c_inputs(m, nrad, nang, lmax, PREC_FP64) for m in mols]
    ms = (_Molecule * n)(*[i[0] for i in ins])
    bs = (_Basis * n)(*[i[1] for i in ins])
    sp = _ScfParams(threshold, min_iter, max_iter, mixing, density_step, 0)
    res = (_ScfResult * n)()
    it = C.c_int()
    _check(lib().dg_batch_scf_run(n, ms, bs, C.byref(ins[0][2]), C.byref(sp), res, C.byref(it)))
    return [{k: getattr(r, k) for k, _ in _ScfResult._fields_} for r in res], it.value


BASIS_DIR = os.path.join(_HERE, "data", "basis")
HARTREE_BECKE = 0
HARTREE_ERI = 1


class InputFile:
    """A reference input file (molecules/<name>.in) and its basis set, read by
    dg_input_read (src/settings.cpp:30-187, src/molecule.cpp:49-238).

    Attributes: ``name``, ``basis``, ``units``, ``Z`` and ``xyz`` (bohr),
    ``radial_points``, ``angular_points``, ``lebedev_order`` (the reference's
    index), ``lmax``, ``hartree`` (HARTREE_BECKE or HARTREE_ERI).  It stands in
    for a :class:`dftcxx_amd.molecule.Molecule` wherever one is taken: its
    ``shells()`` are the basis file's shells in the reference's CGF order.
    """

    def __init__(self, path, basis_dir=None):
        self._h = None
        h = C.c_void_p()
        _check(lib().dg_input_read(os.fsencode(path), os.fsencode(basis_dir or BASIS_DIR), C.byref(h)))
        self._h = h
        m, b, g = _Molecule(), _Basis(), _GridParams()
        li, he = C.c_int(), C.c_int()
        _check(lib().dg_input_molecule(h, C.byref(m)))
        _check(lib().dg_input_basis(h, C.byref(b)))
        _check(lib().dg_input_grid(h, C.byref(g), C.byref(li), C.byref(he)))
        self.Z = np.array(m.Z[:m.natoms], dtype=np.int32)
        self.xyz = np.array(m.xyz[:3 * m.natoms], dtype=np.float64).reshape(-1, 3)
        nsh = b.nshell
        nprim = np.array(b.nprim[:nsh], dtype=np.int32)
        offset = np.array(b.offset[:nsh], dtype=np.int32)
        ntot = int((offset + nprim).max()) if nsh else 0
        L = np.array(b.L[:nsh], dtype=np.int32)
        self._shells = {"atom": np.array(b.atom[:nsh], dtype=np.int32), "L": L, "nprim": nprim,
                        "offset": offset, "alpha": np.array(b.alpha[:ntot], dtype=np.float64),
                        "coef": np.array(b.coef[:ntot], dtype=np.float64),
                        "nbf": int(sum((l + 1) * (l + 2) // 2 for l in L))}
        self.radial_points, self.angular_points, self.lmax = g.n_radial, g.n_angular, g.lmax
        self.lebedev_order, self.hartree = li.value, he.value
        self.name = self.value("name")
        self.basis = self.value("basis")
        self.units = self.value("units") if self.has("units") else ""

    def value(self, key):
        """Settings::get_value: the trimmed value of a key (DgError "Could not find <key>")."""
        v = C.c_char_p()
        _check(lib().dg_input_value(self._h, key.encode(), C.byref(v)))
        return v.value.decode()

    def has(self, key):
        v = C.c_char_p()
        return lib().dg_input_value(self._h, key.encode(), C.byref(v)) == 0

    def close(self):
        if self._h is not None:
            lib().dg_input_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def natoms(self):
        return len(self.Z)

    @property
    def nelec(self):
        return int(self.Z.sum())

    def shells(self):
        return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self._shells.items()}

    def nbf(self):
        return self._shells["nbf"]


def read_input(path, basis_dir=None):
    """Parse a reference input file; see :class:`InputFile`."""
    return InputFile(path, basis_dir)


class Eri:
    """Analytic two-electron integrals of a basis on the GPU (dg_eri_*): every unique
    (ij|kl) once (src/dft.cpp:254-288), J = sum_kl P_kl (ij|kl) from them
    (src/dft.cpp:452-477)."""

    def __init__(self, mol):
        self.e = None
        m, b, g, keep = _c_inputs(mol, 5, 6, 0, PREC_FP64)
        h = C.c_void_p()
        _check(lib().dg_eri_create(C.byref(m), C.byref(b), C.byref(h)))
        self.e = h
        n, nb = C.c_long(), C.c_int()
        _check(lib().dg_eri_count(self.e, C.byref(n), C.byref(nb)))
        self.nints, self.nbf = n.value, nb.value

    def close(self):
        if self.e is not None:
            lib().dg_eri_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def integrals(self):
        """Packed unique integrals in the reference's teindex order."""
        out = np.empty(self.nints)
        _check(lib().dg_eri_get(self.e, _d(out)))
        return out

    def coulomb(self, P):
        Pc = _f64(np.asarray(P).T)
        J = np.empty((self.nbf, self.nbf))
        _check(lib().dg_eri_coulomb(self.e, _d(Pc), _d(J)))
        return J.T.copy()


def teindex(i, j, k, l):
    """Integrator::teindex (src/integrals.cpp:470-486)."""
    if i < j:
        i, j = j, i
    if k < l:
        k, l = l, k
    ij, kl = i * (i + 1) // 2 + j, k * (k + 1) // 2 + l
    if ij < kl:
        ij, kl = kl, ij
    return ij * (ij + 1) // 2 + kl
