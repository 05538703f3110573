"""GPU checks of the batched build (dg_batch_*): every molecule of a batch
against its own single-molecule handle, to 1e-12 relative, graph replays,
and the autotuner.  The single-molecule path is itself checked against the
oracle in test_gpu_parity.py."""
import numpy as np
import pytest

from dftcxx_amd import gridlib as G
from dftcxx_amd import molecule as M

from systems import core_guess

pytestmark = pytest.mark.gpu


def _guess(mol):
    S, T, V = G.one_electron(mol)
    return core_guess(mol, S, T, V)


def _single(mol, grid, P):
    g = G.MolecularGrid(mol, *grid)
    g.set_density(P)
    nbefore = g.correct_densities()
    J = g.calculate_hartree_potential()
    XC, exc = g.calculate_exchange_correlation()
    return J, XC, exc, nbefore


def _compare(batch, mols, grid, Ps, J, XC, sc, which):
    for m in which:
        Js, XCs, exc, nb = _single(mols[m], grid, Ps[m])
        tag = f"molecule {m} ({mols[m].name}/{mols[m].basis})"
        assert np.abs(J[m] - Js).max() < 1e-12 * np.abs(Js).max(), tag
        assert np.abs(XC[m] - XCs).max() < 1e-12 * np.abs(XCs).max(), tag
        assert abs(sc[m, 0] - exc) < 1e-12 * abs(exc), tag
        assert abs(sc[m, 1] - nb) < 1e-12 * nb, tag
        assert abs(sc[m, 2] - mols[m].nelec) < 1e-11, tag
        ej = 2.0 * np.sum(Ps[m] * Js)
        assert abs(sc[m, 3] - ej) < 1e-11 * abs(ej), tag


def test_batch_matches_single_molecule_handles():
    mols = M.batch(count=12, seed=7)
    grid = (20, 50, 6)
    print("[case] batch of 12 at 20x50", flush=True)
    b = G.Batch(mols, *grid)
    assert b.nmol == 12 and list(b.nbf) == [m.nbf() for m in mols]
    Ps = [_guess(m) for m in mols]
    J, XC, sc = b.build(Ps)
    _compare(b, mols, grid, Ps, J, XC, sc, range(12))
    # the Becke weights are those of the single handles bit for bit; the densities
    # to rounding (the renormalisation sums its cells in a different order)
    for m in (0, 5, 11):
        g = G.MolecularGrid(mols[m], *grid)
        g.set_density(Ps[m])
        g.correct_densities()
        sl = slice(b.point0[m], b.point0[m + 1])
        assert np.array_equal(b.get_weights()[sl], g.get_weights())
        r = g.get_densities()
        assert np.abs(b.get_densities()[sl] - r).max() < 1e-14 * r.max()


def test_batch_matches_oracle():
    # batched molecules straight against the oracle (not only through the single handle)
    from oracle import oracle as O
    mols = M.batch(count=12, seed=7)
    grid = (20, 50, 6)
    print("[case] batch of 12 at 20x50 vs oracle", flush=True)
    b = G.Batch(mols, *grid)
    Ps = [_guess(m) for m in mols]
    J, XC, sc = b.build(Ps)
    for m in (0, 4, 9, 11):
        o = O.Grid(mols[m], *grid)
        o.set_density(Ps[m])
        o.correct_densities()
        Jo = o.hartree()
        XCo, exco = o.xc()
        tag = f"molecule {m} ({mols[m].name}/{mols[m].basis})"
        assert np.abs(J[m] - Jo).max() < 1e-10 * np.abs(Jo).max(), tag
        assert np.abs(XC[m] - XCo).max() < 1e-10 * np.abs(XCo).max(), tag
        assert abs(sc[m, 0] - exco) < 1e-10 * abs(exco), tag
        assert abs(sc[m, 2] - mols[m].nelec) < 1e-11, tag


def test_batch_graph_replays_and_autotune():
    mols = M.batch(count=6, seed=11)
    grid = (16, 26, 4)
    print("[case] batch graph replays", flush=True)
    b = G.Batch(mols, *grid)
    P1 = [_guess(m) for m in mols]
    P2 = [0.9 * P for P in P1]
    r1 = b.build(P1)
    r2 = b.build(P2)
    for _ in range(3):  # replays of the same graph, with eager host work between
        _ = [np.linalg.eigh(P) for P in P1]
        a = b.build(P1)
        c = b.build(P2)
        for x, y in ((a, r1), (c, r2)):
            assert all(np.array_equal(u, v) for u, v in zip(x[0], y[0]))
            assert all(np.array_equal(u, v) for u, v in zip(x[1], y[1]))
            assert np.array_equal(x[2], y[2])
    chunk, item, ms = b.autotune()
    assert chunk in (256, 512, 1024, 2048, 4096) and item in (256, 512, 1024, 2048, 4096)
    assert np.all(ms > 0)
    t = b.build(P1)
    for u, v in zip(t[0] + t[1], r1[0] + r1[1]):
        assert np.abs(u - v).max() < 1e-13 * np.abs(v).max()
    assert np.abs(t[2] - r1[2]).max() < 1e-12 * np.abs(r1[2]).max()


def test_batch_of_256_full_grid():
    # 256 seeded geometries of H2O, CH4, N2, CO, ethane (STO-3G and 6-31G),
    # 75 x 302, lmax 8: sampled molecules against single handles
    mols = M.batch()
    grid = (75, 302, 8)
    print("[case] batch of 256 at 75x302", flush=True)
    b = G.Batch(mols, *grid)
    Ps = [_guess(m) for m in mols]
    J, XC, sc = b.build(Ps)
    assert np.all(np.abs(sc[:, 2] - [m.nelec for m in mols]) < 1e-10)
    _compare(b, mols, grid, Ps, J, XC, sc, range(0, 256, 37))


def test_batch_rejects_single_molecule_calls():
    mols = M.batch(count=2, seed=3)
    b = G.Batch(mols, 10, 26, 4)
    with pytest.raises(G.DgError):
        G._check(G.lib().dg_set_density(b.h, G._d(np.zeros(4)), 2))
    # after a batch build the handle holds a density; its single-molecule calls
    # still return DG_ERR_STATE before any launch
    b.build([_guess(m) for m in mols])
    L = G.lib()
    out = np.zeros(64 * 64)
    x = np.zeros(1)
    for rc in (L.dg_hartree(b.h, G._d(out)), L.dg_calculate_density(b.h, G._d(x)),
               L.dg_xc(b.h, G._d(out), G._d(x)), L.dg_correct_densities(b.h, G._d(x)),
               L.dg_build_device(b.h, 1, 0, 0, 0)):
        assert rc == 4 and "batch handle" in L.dg_last_error().decode()


@pytest.mark.parametrize("step", [G.STEP_HOST, G.STEP_DEVICE], ids=["host_step", "device_step"])
def test_batch_scf_matches_single_scf(step):
    mols = M.batch(count=6, seed=21)
    grid = (30, 50, 6)
    print(f"[case] batch SCF of 6 at 30x50, step {step}", flush=True)
    res, total = G.batch_scf_run(mols, *grid, threshold=1e-6, density_step=step)
    assert res[0]["device_step"] == (1 if step == G.STEP_DEVICE else 0)
    assert total == max(r["iterations"] for r in res)
    for m, r in zip(mols, res):
        _, ref, _ = G.scf_run(m, *grid, threshold=1e-6)
        assert r["iterations"] == ref["iterations"], m.name
        assert abs(r["energy"] - ref["energy"]) < 1e-9, (m.name, r["energy"], ref["energy"])
        assert abs(r["nelec"] - m.nelec) < 1e-10
