"""Host-side checks that need no GPU: the C-ABI library, host integrals and
eigen-solver, the molecule/basis builders and the workload generators."""
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from dftcxx_amd import gridlib as G
from dftcxx_amd import lebedev
from dftcxx_amd import molecule as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BASIS = "/root/reference/basis"


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "dftgrid.h")).read()
    declared = set(re.findall(r"^\w[\w\s\*]*?\b(dg_\w+)\(", hdr, flags=re.M))
    assert declared == set(G.EXPORTS)
    L = G.lib()
    for name in declared:
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", G.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (dg_\w+)", out))
    assert declared <= exported


def test_library_has_gfx950_code_object():
    # the library built here carries a code object for gfx950 (offload target triple)
    with open(G.LIB_PATH, "rb") as f:
        data = f.read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_lebedev_availability():
    L = G.lib()
    for n in lebedev.available():
        assert L.dg_lebedev_available(n) == 1
    assert L.dg_lebedev_available(7) == 0
    assert L.dg_version() >= 100


def test_reference_enum_mapping():
    assert [lebedev.points_from_reference_index(i) for i in range(11)] == \
        [6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194]
    with pytest.raises(ValueError):
        lebedev.points_from_reference_index(11)
    g = G.MolecularGrid(M.h2o())
    g.set_grid_parameters(50, 7, 8)  # the reference's config-1 input
    assert (g.radial_points, g.angular_points, g.lmax) == (50, 110, 8)


# C10H22 (94 shells) takes the threaded path of the host integrals
@pytest.mark.parametrize("make", [M.h2o, M.benzene, lambda: M.ch4("p631"), lambda: M.alkane(3), lambda: M.alkane(10)])
def test_host_integrals_match_oracle(oracle_lib, make):
    mol = make()
    S, T, V = G.one_electron(mol)
    So, To, Vo = oracle_lib.one_electron(mol)
    for a, b in ((S, So), (T, To), (V, Vo)):
        assert np.abs(a - b).max() < 1e-12 * max(1.0, np.abs(b).max())
    assert np.abs(np.diag(S) - 1.0).max() < 1e-3  # contracted functions are nearly normalised


def test_host_integrals_d_shell(oracle_lib):
    from systems import with_d_shell
    mol = with_d_shell(M.co("sto3g"))
    S, T, V = G.one_electron(mol)
    So, To, Vo = oracle_lib.one_electron(mol)
    for a, b in ((S, So), (T, To), (V, Vo)):
        assert np.abs(a - b).max() < 1e-12 * max(1.0, np.abs(b).max())


def test_host_eigh():
    rng = np.random.default_rng(2)
    for n in (1, 2, 7, 66):
        A = rng.standard_normal((n, n))
        A = A + A.T
        w, Z = G.eigh(A)
        assert np.all(np.diff(w) >= 0)
        assert np.abs(w - np.linalg.eigvalsh(A)).max() < 1e-12 * max(1, np.abs(w).max())
        assert np.abs(Z.T @ Z - np.eye(n)).max() < 1e-12
        assert np.abs(A @ Z - Z * w).max() < 1e-11 * max(1, np.abs(w).max())
    # degenerate spectrum
    Q, _ = np.linalg.qr(rng.standard_normal((6, 6)))
    A = Q @ np.diag([1, 1, 1, 2, 2, 5.0]) @ Q.T
    w, Z = G.eigh(A)
    assert np.abs(w - [1, 1, 1, 2, 2, 5]).max() < 1e-13


def test_error_reporting():
    mol = M.h2o()
    s = mol.shells()
    bad = dict(s)
    bad["L"] = s["L"].copy()
    bad["L"][0] = 3
    keep = [np.ascontiguousarray(mol.Z, np.int32), np.ascontiguousarray(mol.xyz)]
    import ctypes as C
    m = G._Molecule(mol.natoms, G._i(keep[0]), G._d(keep[1]))
    arrs = [np.ascontiguousarray(bad[k], np.int32) for k in ("atom", "L", "nprim", "offset")]
    fl = [np.ascontiguousarray(bad[k]) for k in ("alpha", "coef")]
    b = G._Basis(len(arrs[0]), *[G._i(a) for a in arrs], *[G._d(a) for a in fl])
    n = mol.nbf()
    out = np.empty((n, n))
    rc = G.lib().dg_one_electron(C.byref(m), C.byref(b), G._d(out), None, None)
    assert rc == 1
    assert "L must be" in G.lib().dg_last_error().decode()
    with pytest.raises(G.DgError):
        G._check(G.lib().dg_eigh(0, G._d(out), G._d(out), G._d(out)))


# --- molecules, basis data and workloads ---------------------------------------------

def test_workload_sizes():
    # N_bf and N_pts of the four single-molecule workloads
    expect = {"h2o_sto3g": (7, 16500), "benzene_631g": (66, 271800),
              "c10h22_631g": (134, 724800), "c30h62_631g": (394, 2083800)}
    for name, (make, nrad, nang) in M.WORKLOADS.items():
        mol = make()
        assert (mol.nbf(), mol.natoms * nrad * nang) == expect[name]


def test_cgf_order_is_by_element_then_atom():
    mol = M.benzene()
    s = mol.shells()
    # H (Z=1) shells come first although the carbons are listed first
    assert list(s["atom"][:12]) == [a for a in range(6, 12) for _ in range(2)]
    assert all(mol.Z[a] == 6 for a in s["atom"][12:])
    assert list(s["L"][12:17]) == [0, 0, 1, 0, 1]


def test_geometries():
    A = M.ANGSTROM_TO_BOHR
    h2o = M.h2o()
    d = np.linalg.norm(h2o.xyz[1:] - h2o.xyz[0], axis=1) / A
    assert np.allclose(d, 0.9572, atol=1e-12)
    v1, v2 = h2o.xyz[1] - h2o.xyz[0], h2o.xyz[2] - h2o.xyz[0]
    ang = math.degrees(math.acos(v1 @ v2 / np.linalg.norm(v1) / np.linalg.norm(v2)))
    assert abs(ang - 104.52) < 1e-9
    for n in (10, 30):
        mol = M.alkane(n)
        C = mol.xyz[:n] / A
        assert np.allclose(np.linalg.norm(np.diff(C, axis=0), axis=1), 1.54, atol=1e-12)
        H = mol.xyz[n:] / A
        dmin = np.linalg.norm(H[:, None] - C[None], axis=-1).min(1)
        assert np.allclose(dmin, 1.09, atol=1e-12)
        assert mol.natoms == 3 * n + 2
        # every carbon has four neighbours at bonding distance
        allx = mol.xyz / A
        dd = np.linalg.norm(allx[:n, None] - allx[None], axis=-1)
        assert np.all(((dd > 0.5) & (dd < 1.6)).sum(1) == 4)
        # no two atoms closer than 1.7 A apart from bonds
        dall = np.linalg.norm(allx[:, None] - allx[None], axis=-1) + 10 * np.eye(len(allx))
        assert dall.min() > 1.0


def test_splitmix64_known_answer():
    r = M.SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                0x06C45D188009454F]


def test_batch_is_seeded_and_mixed():
    b1, b2 = M.batch(), M.batch()
    assert len(b1) == 256
    assert all(np.array_equal(x.xyz, y.xyz) and x.basis == y.basis for x, y in zip(b1, b2))
    names = {m.name for m in b1}
    assert names == {"h2o", "ch4", "n2", "co", "ethane"}
    assert {m.basis for m in b1} == {"sto3g", "p631"}
    assert 2 <= min(m.natoms for m in b1) and max(m.natoms for m in b1) <= 8


def test_basis_json_matches_basis_files():
    """The JSON basis tables equal the shipped .dat files (the reference's basis/*.dat),
    which are byte-identical to the reference's when its tree is mounted."""
    from dftcxx_amd.data.make_basis import parse_dat
    shipped = os.path.join(ROOT, "dftcxx_amd", "data", "basis")
    for name in ("sto3g", "sto6g", "p321", "p631"):
        path = os.path.join(shipped, name + ".dat")
        if os.path.isdir(REF_BASIS):
            with open(path, "rb") as a, open(os.path.join(REF_BASIS, name + ".dat"), "rb") as b:
                assert a.read() == b.read()
        ref = parse_dat(path)
        mine = M.load_basis(name)
        for z in range(1, 19):
            assert [(M.SHELL_L[k], [p[0] for p in pr], [p[1] for p in pr]) for k, pr in ref[z]] == \
                [(l, list(a), list(c)) for l, a, c in mine[z]]


def test_bench_traffic_key_picks_the_fp64_two_output_kernels():
    # roofline.traffic must come from the kernel the build's timing is for: the
    # two-output FP64 contraction, not the single-output one the host path runs
    sys.path.insert(0, ROOT)
    from bench import traffic_key
    names = ["contract2d_lds<1, 4, 3, true>", "contract2d_lds<2, 2, 4, true>", "contract2d_lds_f32<2, 2, 12, true>",
             "density<25, 4, 1>", "hartree_cross<8, 2>", "vreduce"]
    assert traffic_key("contract", names) == "contract2d_lds<2, 2, 4, true>"
    assert traffic_key("density", names) == "density<25, 4, 1>"
    assert traffic_key("hartree_cross", names) == "hartree_cross<8, 2>"
    small = ["contract_small<5, 1, 2>", "contract_small<5, 2, 2>", "density_pl<5>"]
    assert traffic_key("contract", small) == "contract_small<5, 2, 2>"
    assert traffic_key("density", small) == "density_pl<5>"
    assert traffic_key("contract", ["vreduce"]) is None


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_cpp_facade_compiles(tmp_path):
    # the reference-shaped C++ facade is header-only over the C ABI
    src = tmp_path / "use_facade.cpp"
    src.write_text('#include "dftcxx/molecular_grid.hpp"\n'
                   "int main() { dftcxx::MoleculeData m; dftcxx::MolecularGrid g(m);\n"
                   "  std::vector<double> dP, gr; (void)&dftcxx::MolecularGrid::xc_adjoint;\n"
                   "  return dg_version() >= 100 ? 0 : 1; }\n")
    exe = tmp_path / "use_facade"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        str(src), "-o", str(exe), "-L", os.path.dirname(G.LIB_PATH), "-ldftgrid",
                        "-Wl,-rpath," + os.path.dirname(G.LIB_PATH)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert subprocess.run([str(exe)]).returncode == 0


@pytest.mark.parametrize("change,msg", [
    (dict(nrad=4), "n_radial"),
    (dict(nang=7), "Lebedev"),
    (dict(lmax=13), "lmax"),
    (dict(Z=19), "Z outside"),
    (dict(coincide=True), "coincide"),
    (dict(L=3), "L must be"),
    (dict(alpha=-1.0), "exponent"),
    (dict(screen=-1.0), "screen"),
    (dict(precision=7), "precision"),
])
def test_create_rejects_invalid_input(change, msg):
    # argument checks run before any device call, so they hold without a GPU
    import ctypes as C
    mol = M.h2o()
    Z = np.array(mol.Z, np.int32)
    xyz = np.array(mol.xyz)
    if "Z" in change:
        Z[0] = change["Z"]
    if change.get("coincide"):
        xyz[1] = xyz[0]
    s = mol.shells()
    L = np.array(s["L"], np.int32)
    alpha = np.array(s["alpha"])
    if "L" in change:
        L[0] = change["L"]
    if "alpha" in change:
        alpha[0] = change["alpha"]
    keep = [Z, xyz, np.ascontiguousarray(s["atom"], np.int32), L, np.ascontiguousarray(s["nprim"], np.int32),
            np.ascontiguousarray(s["offset"], np.int32), alpha, np.ascontiguousarray(s["coef"])]
    m = G._Molecule(mol.natoms, G._i(keep[0]), G._d(keep[1]))
    b = G._Basis(len(keep[2]), G._i(keep[2]), G._i(keep[3]), G._i(keep[4]), G._i(keep[5]), G._d(keep[6]),
                 G._d(keep[7]))
    prm = G._GridParams(change.get("nrad", 20), change.get("nang", 26), change.get("lmax", 4),
                        change.get("precision", 0), change.get("screen", 0.0))
    h = C.c_void_p()
    rc = G.lib().dg_create(C.byref(m), C.byref(b), C.byref(prm), C.byref(h))
    assert rc == 1 and not h.value
    assert msg in G.lib().dg_last_error().decode()


def test_host_integrals_independent_of_thread_count():
    # the threaded host integrals (C10H22, 94 shells) are bitwise those of one thread
    import os
    import subprocess
    import sys
    code = ("import hashlib, numpy as np; from dftcxx_amd import gridlib as G, molecule as M; "
            "S, T, V = G.one_electron(M.alkane(10)); "
            "print(hashlib.sha256(np.concatenate([S.ravel(), T.ravel(), V.ravel()]).tobytes()).hexdigest())")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = []
    for nt in ("1", "8"):
        env = dict(os.environ, OMP_NUM_THREADS=nt)
        r = subprocess.run([sys.executable, "-c", code], cwd=root, env=env, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr
        out.append(r.stdout.strip())
    assert out[0] == out[1]


@pytest.mark.parametrize("field,value,msg", [
    ("max_iter", 0, "max_iter"),
    ("min_iter", -1, "min_iter"),
    ("threshold", float("nan"), "threshold"),
    ("threshold", -1e-6, "threshold"),
    ("mixing", 1.0, "mixing"),
    ("mixing", -0.1, "mixing"),
    ("density_step", 3, "density_step must be"),
    ("density_step", -1, "density_step must be"),
    ("hartree", 2, "hartree must be"),
])
def test_scf_rejects_invalid_params(field, value, msg):
    # the SCF parameters are checked before any device call, so this holds without a GPU
    import ctypes as C
    mol = M.h2o()
    m, b, g, keep = G._c_inputs(mol, 20, 26, 4, G.PREC_FP64)
    kw = dict(threshold=1e-6, min_iter=3, max_iter=100, mixing=0.5, density_step=0, hartree=0)
    kw[field] = value
    sp = G._ScfParams(kw["threshold"], kw["min_iter"], kw["max_iter"], kw["mixing"], kw["density_step"],
                      kw["hartree"])
    res = G._ScfResult()
    rc = G.lib().dg_scf_run(C.byref(m), C.byref(b), C.byref(g), C.byref(sp), C.byref(res), None, None)
    assert rc == 1 and msg in G.lib().dg_last_error().decode()
    rs = (G._ScfResult * 1)()
    ms = (G._Molecule * 1)(m)
    bs = (G._Basis * 1)(b)
    rc = G.lib().dg_batch_scf_run(1, ms, bs, C.byref(g), C.byref(sp), rs, None)
    assert rc == 1 and msg in G.lib().dg_last_error().decode()
