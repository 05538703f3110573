"""Benchmark of one SCF iteration's grid build on one MI355X.

A step is one grid build on synthetic input already resident in HBM: from the
density matrix P to rho, its renormalisation, Slater+VWN, the Becke-Poisson
Hartree potential, J = Phi^T diag(w V/2) Phi and XC = Phi^T diag(w v_xc) Phi
(the reference's per-iteration grid work, src/dft.cpp:330-433).  The metric is
N_pts * N_bf^2 per second.  Default workload: linear C30H62, 6-31G (394 CGFs),
92 atoms x 75 radial x 302 Lebedev points (2,083,800 points), lmax 8; P is the
core-Hamiltonian guess.

Besides the throughput line it reports, for the kernel with the largest share
of the build, its achieved rate against the MI355X peak (timed with HIP events
on the stream the kernel runs on), the screened and reduced-precision builds,
SCF energies and s/iteration, the adjoint, and the reference-faithful CPU
restatement timed on this host (on the whole build for grids up to 300k
points, else on the points of one atomic cell with the per-stage times scaled
to the whole grid).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
HARTREE_FLOPS_PER_TERM = 10.0


def traffic_key(label, names):
    """The rocprofv3 kernel name (with template arguments) that runs build stage
    `label` in an FP64 dg_build_device, among `names`; None if absent."""
    for k in sorted(names):
        base = k.split("<")[0]
        args = [a.strip() for a in k[k.find("<") + 1:k.rfind(">")].split(",")] if "<" in k else []
        # two-output (J and XC) FP64 variants: contract2d_lds<NOUT, ...>,
        # contract_small<NT, NOUT, ...>
        if label == "contract" and ((base == "contract2d_lds" and args[:1] == ["2"]) or
                                    (base == "contract_small" and len(args) > 1 and args[1] == "2")):
            return k
        if label == "density" and base in ("density", "density_pl"):
            return k
        if label == "hartree_cross" and base == "hartree_cross":
            return k
    return None
CPU_SAMPLE_MAX_POINTS = 300000


def kernel_work(name, mol_info):
    """(algorithmic flops of one launch, the unit that bounds it) of a build kernel,
    or (None, None) for kernels without a flop model.

    One convention throughout: the symmetric products count only the pairs
    i <= j (the work the result needs), 2 flops per multiply-add; the Hartree
    cross term counts 10 flops per (point, atom, lm) spline + harmonic term."""
    npts, nbf, nat, nlm = mol_info["npts"], mol_info["nbf"], mol_info["natoms"], mol_info["nlm"]
    if name == "hartree_cross":  # FP64 vector ALU: the kernel issues no MFMA
        return HARTREE_FLOPS_PER_TERM * npts * (nat - 1) * nlm / mol_info["cross_launches"], "valu_fp64"
    if name == "contract":  # J and XC, FP64 (or FP32 in mixed mode) MFMA
        return 2 * 2.0 * npts * nbf * (nbf + 1) / 2, "mfma"
    if name == "density":
        return 2.0 * npts * nbf * (nbf + 1) / 2 + 2.0 * npts * nbf, "mfma"
    return None, None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mol, nrad, nang, lmax, P, units):
    """Time the oracle (reference-faithful FP64 CPU restatement) on this host.

    Grids up to CPU_SAMPLE_MAX_POINTS are timed whole.  Larger ones are timed on
    the points of the first atomic cell (all atoms still enter the Becke weights,
    Poisson solves and cross-atom potential) and each stage is scaled to the
    whole grid by how the reference's loops run: the density loop is OpenMP over
    cells (rounds of `threads` cells), the XC loop is serial over all points,
    the projection, cross-atom potential and J scale with the number of cells,
    and the per-atom Poisson solves are already done for every atom."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    nat = mol.natoms
    whole = nat * nrad * nang <= CPU_SAMPLE_MAX_POINTS
    ncell = nat if whole else 1
    g = O.Grid(mol, nrad, nang, lmax, ncell=ncell)
    t = {}
    t0 = time.perf_counter()
    g.set_density(P)
    t1 = time.perf_counter()
    g.correct_densities()
    t2 = time.perf_counter()
    g.hartree()
    t3 = time.perf_counter()
    g.xc()
    t4 = time.perf_counter()
    th = g.hartree_times()
    t = {"density": t1 - t0, "correct": t2 - t1, "rho_lm": th[0], "poisson_vcell_splines": th[1],
         "cross_potential": th[2], "J": th[3], "xc": t4 - t3}
    if whole:
        total = t4 - t0
        sample = f"one full grid build of {nat} cells ({total:.1f} s)"
        scaled = None
    else:
        f = nat / ncell
        rounds = -(-nat // threads)
        scaled = {"density": t["density"] * rounds, "correct": t["correct"] * f, "rho_lm": th[0] * f,
                  "poisson_vcell_splines": th[1], "cross_potential": th[2] * f, "J": th[3] * f,
                  "xc": t["xc"] * f}
        total = sum(scaled.values())
        sample = (f"per-iteration build on the {nrad * nang} points of {ncell} of {nat} cells "
                  f"({t4 - t0:.1f} s), stage times scaled to the whole grid ({total:.0f} s)")
    out = {"value": units / total, "unit": "pts*bf^2/s", "cores": threads, "kind": "port",
           "sample": sample, "cpu_model": cpu_model(), "seconds_per_build": total,
           "sample_stage_seconds": t}
    if scaled is not None:
        out["scaled_stage_seconds"] = scaled
    return out


def bench_batch(args, G, M, np, torch):
    """256 seeded small molecules (STO-3G / 6-31G), 75 x 302, one graph launch per step."""
    mols = M.batch()
    nrad, nang, lmax = 75, 302, 8
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    b = G.Batch(mols, nrad, nang, lmax)
    setup = time.perf_counter() - t0
    Ps = []
    for m in mols:
        S, T, V = G.one_electron(m)
        w, U = np.linalg.eigh(S)
        X = U / np.sqrt(w)
        _, Cc = np.linalg.eigh(X.T @ (T + V) @ X)
        C = X @ Cc
        Ps.append(C[:, :m.nelec // 2] @ C[:, :m.nelec // 2].T)
    stream = torch.cuda.Stream(device=dev)
    b.set_stream(stream.cuda_stream)
    dP = torch.tensor(b.pack(Ps), dtype=torch.float64, device=dev)
    dJ = torch.empty(b.packed, dtype=torch.float64, device=dev)
    dX = torch.empty(b.packed, dtype=torch.float64, device=dev)
    dS = torch.empty(4 * b.nmol, dtype=torch.float64, device=dev)
    ptrs = (dP.data_ptr(), dJ.data_ptr(), dX.data_ptr(), dS.data_ptr())
    torch.cuda.synchronize()
    b.build_device(*ptrs)
    torch.cuda.synchronize()
    tuned = b.autotune()
    for _ in range(args.warmup):
        b.build_device(*ptrs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.build_device(*ptrs)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    units = float(sum(int(npt) * int(n) ** 2 for npt, n in zip(np.diff(b.point0), b.nbf)))
    out = {
        "metric": "grid V_xc+J build throughput (N_pts*N_bf^2 per second)",
        "value": units * args.steps / wall, "unit": "pts*bf^2/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded geometries, core-Hamiltonian guess P)",
        "config": {"workload": "batch256", "molecules": b.nmol, "n_pts": b.npts, "n_radial": nrad,
                   "n_angular": nang, "lmax": lmax, "graph": True},
        "grid_setup_s": setup,
        "autotune": {"contract_chunk": tuned[0], "cross_item": tuned[1], "ms": list(map(float, tuned[2]))},
    }
    if not args.no_profile:
        b.profile_enable(True)
        b.profile_reset()
        for _ in range(3):
            b.build_device(*ptrs)
        torch.cuda.synchronize()
        prof = b.profile_read()
        b.profile_enable(False)
        out["kernels_ms_per_step"] = {k: ms / 3 for k, (ms, _) in sorted(prof.items())}
    if not args.no_scf:
        res, its = G.batch_scf_run(mols, nrad, nang, lmax, threshold=1e-6)
        conv = [r["converged"] == 1 for r in res]
        its_conv = [r["iterations"] for r, c in zip(res, conv) if c]
        # s_per_batch_iter: wall time of one batch iteration (every molecule advances)
        out["scf"] = {"batch_iterations": its, "s_per_batch_iter": res[0]["seconds_iter"],
                      "setup_s": res[0]["seconds_setup"], "converged": int(sum(conv)),
                      "not_converged": [{"molecule": m, "name": mols[m].name, "basis": mols[m].basis,
                                         "iterations": res[m]["iterations"], "last_dE": res[m]["delta"]}
                                        for m in range(len(res)) if not conv[m]],
                      "max_converged_iterations": max(its_conv) if its_conv else None,
                      "s_until_last_converged": (max(its_conv) * res[0]["seconds_iter"]) if its_conv else None}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c30h62_631g")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--screen", type=float, default=0.0,
                    help="amplitude screening threshold of the timed build (0 = dense, as the reference)")
    ap.add_argument("--variant-screen", type=float, default=1e-10,
                    help="screening threshold of the screened comparison builds")
    ap.add_argument("--no-mixed", action="store_true", help="skip the reduced-precision comparison")
    ap.add_argument("--no-scf", action="store_true", help="skip the SCF energy / s-per-iteration runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip timing dE_xc/dP and dE_xc/dR")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (HBM-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is designed, not built.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_calibration:
        # 256 MiB device-to-device copies: the HBM counters of this known traffic
        # (256 MiB read + 256 MiB written per copy) scale the kernel counters
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 8, "copies": 5}))
        return

    from dftcxx_amd import gridlib as G
    from dftcxx_amd import molecule as M

    if args.workload == "batch256":
        return bench_batch(args, G, M, np, torch)
    make, nrad, nang = M.WORKLOADS[args.workload]
    lmax = 8
    mol = make()
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")

    # synthetic input: core-Hamiltonian guess density matrix
    S, T, V = G.one_electron(mol)
    w, U = np.linalg.eigh(S)
    X = U / np.sqrt(w)
    _, Cc = np.linalg.eigh(X.T @ (T + V) @ X)
    C = X @ Cc
    P = C[:, :mol.nelec // 2] @ C[:, :mol.nelec // 2].T

    def timed_builds(precision, screen):
        """(grid, ptrs, wall seconds, device ms) for warmup + steps builds on one stream."""
        t0 = time.perf_counter()
        grid = G.MolecularGrid(mol, nrad, nang, lmax, precision=precision, screen=screen)
        setup = time.perf_counter() - t0
        nbf = grid.nbf
        stream = torch.cuda.Stream(device=dev)
        grid.set_stream(stream.cuda_stream)
        dP = torch.tensor(P.T.copy(), dtype=torch.float64, device=dev)
        dJ = torch.empty((nbf, nbf), dtype=torch.float64, device=dev)
        dX = torch.empty((nbf, nbf), dtype=torch.float64, device=dev)
        dS = torch.empty(4, dtype=torch.float64, device=dev)
        ptrs = (dP.data_ptr(), dJ.data_ptr(), dX.data_ptr(), dS.data_ptr())
        torch.cuda.synchronize()
        for _ in range(args.warmup):
            grid.build_device(*ptrs)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            grid.build_device(*ptrs)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return grid, ptrs, (dP, dJ, dX, dS), wall, e0.elapsed_time(e1), setup

    prec = G.PREC_MIXED if args.precision == "mixed" else G.PREC_FP64
    grid, ptrs, bufs, wall, dev_ms, setup_s = timed_builds(prec, args.screen)
    nbf, npts = grid.nbf, grid.npts
    ms_per_step = 1e3 * wall / args.steps
    units = float(npts) * nbf * nbf
    value = units * args.steps / wall
    scal = bufs[3].cpu().numpy()

    out = {
        "metric": "grid V_xc+J build throughput (N_pts*N_bf^2 per second)",
        "value": value,
        "unit": "pts*bf^2/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == G.PREC_FP64 else "f64+f32",
        "data": "synthetic (core-Hamiltonian guess P)",
        "config": {"workload": args.workload, "molecule": mol.name, "basis": mol.basis,
                   "n_bf": nbf, "n_atoms": mol.natoms, "n_radial": nrad, "n_angular": nang,
                   "lmax": lmax, "n_pts": npts, "precision": args.precision, "screen": args.screen},
        "device_ms_per_step": dev_ms / args.steps,
        "grid_setup_s": setup_s,
        "e_xc": float(scal[0]),
        "n_elec": float(scal[2]),
    }

    if not args.no_profile:
        grid.profile_enable(True)
        grid.profile_reset()
        nprof = max(3, min(args.steps, 10))
        for _ in range(nprof):
            grid.build_device(*ptrs)
        torch.cuda.synchronize()
        prof = grid.profile_read()
        grid.profile_enable(False)
        total = sum(ms for ms, _ in prof.values())
        out["kernels_ms_per_step"] = {k: ms / nprof for k, (ms, _) in sorted(prof.items())}
        info = {"npts": npts, "nbf": nbf, "natoms": mol.natoms, "nlm": (lmax + 1) ** 2,
                "cross_launches": max(1, prof.get("hartree_cross", (0, nprof))[1] // nprof)}
        dom = max(prof.items(), key=lambda kv: kv[1][0])[0]
        ms, cnt = prof[dom]
        avg_s = ms / cnt / 1e3
        work, bound = kernel_work(dom, info)
        roof = {"kernel": dom, "share_of_build": prof[dom][0] / total if total else None}
        if work is not None:
            ach = work / avg_s / 1e12
            peak = PEAK_FP32_MFMA_TFLOPS if (dom == "contract" and prec == G.PREC_MIXED) else PEAK_FP64_TFLOPS
            roof.update({"bound": bound, "achieved": ach, "peak": peak,
                         "unit": "TFLOP/s", "frac": ach / peak, "flops_per_launch": work})
        else:
            roof.update({"bound": "hbm", "achieved": None, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": None})
        roof["avg_launch_ms"] = avg_s * 1e3
        # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
        # summary of this workload (keyed by kernel name)
        roof["traffic"] = None
        tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
        if args.precision != "fp64" or args.screen != 0.0:
            tpath = ""
        if os.path.exists(tpath):
            with open(tpath) as f:
                t = json.load(f)
            kname = traffic_key(dom, t.keys())
            if kname is not None:
                roof["traffic"] = t[kname]
                roof["traffic_source"] = f"profiles/traffic_{args.workload}.json[{kname}]"
        out["roofline"] = roof

    # the reference's call sequence with host buffers (P in, J and XC out over PCIe,
    # each call synchronous): a PCIe-inclusive rate, reported beside the value
    for _ in range(2):
        grid.set_density(P)
        grid.correct_densities()
        grid.calculate_hartree_potential()
        grid.calculate_exchange_correlation()
    nhost = 10
    t0 = time.perf_counter()
    for _ in range(nhost):
        grid.set_density(P)
        grid.correct_densities()
        grid.calculate_hartree_potential()
        grid.calculate_exchange_correlation()
    host_s = (time.perf_counter() - t0) / nhost
    out["host_buffer_path"] = {"ms_per_build": 1e3 * host_s, "value": units / host_s, "unit": "pts*bf^2/s",
                               "calls": "set_density, correct_densities, calculate_hartree_potential, "
                                        "calculate_exchange_correlation"}

    if not args.no_adjoint:
        # dE_xc/dP and dE_xc/dR of the build above (c varying), second call timed
        grid.build_device(*ptrs)
        torch.cuda.synchronize()
        grid.xc_adjoint(hold_c=False)
        grid.profile_enable(True)
        grid.profile_reset()
        t0 = time.perf_counter()
        _, grad = grid.xc_adjoint(hold_c=False)
        adj_s = time.perf_counter() - t0
        prof = grid.profile_read()
        grid.profile_enable(False)
        out["adjoint"] = {"ms": 1e3 * adj_s,
                          "kernels_ms": {k: ms for k, (ms, _) in sorted(prof.items())},
                          "max_abs_grad": float(np.abs(grad).max()),
                          "max_abs_sum_grad": float(np.abs(grad.sum(0)).max())}

    if not args.no_mixed and prec == G.PREC_FP64 and args.screen == 0.0:
        wall_screened = None
        # the same build screened in FP64, and with FP32-MFMA Fock contractions (screened and not)
        J6, X6 = bufs[1].cpu().numpy(), bufs[2].cpu().numpy()
        out["variants"] = {}
        for name, pr, sc in (("fp64_screened", G.PREC_FP64, args.variant_screen),
                             ("mixed", G.PREC_MIXED, 0.0),
                             ("mixed_screened", G.PREC_MIXED, args.variant_screen)):
            gv, pv, bv, wv, dv, _ = timed_builds(pr, sc)
            if name == "fp64_screened":
                wall_screened = wv
            Jv, Xv = bv[1].cpu().numpy(), bv[2].cpu().numpy()
            out["variants"][name] = {
                "precision": "fp64" if pr == G.PREC_FP64 else "mixed", "screen": sc,
                "value": units * args.steps / wv, "ms_per_step": 1e3 * wv / args.steps,
                "speedup_vs_dense_fp64": wall / wv,
                "speedup_vs_screened_fp64": wall_screened / wv if wall_screened else None,
                "J_max_rel_err": float(np.abs(Jv - J6).max() / np.abs(J6).max()),
                "XC_max_rel_err": float(np.abs(Xv - X6).max() / np.abs(X6).max()),
                "e_xc_err": float(bv[3].cpu().numpy()[0] - scal[0]),
            }
            if not args.no_profile:
                gv.profile_enable(True)
                gv.profile_reset()
                for _ in range(3):
                    gv.build_device(*pv)
                torch.cuda.synchronize()
                out["variants"][name]["kernels_ms_per_step"] = {
                    k: ms / 3 for k, (ms, _) in sorted(gv.profile_read().items())}
            del gv, bv

    if not args.no_scf:
        # the reference SCF (plain 0.5 damping, |dE| < 1e-6, at most 100 iterations);
        # density_step AUTO forms P on the device for 128 functions and more
        scf = {}
        for name, pr, sc in (("fp64", G.PREC_FP64, 0.0), ("fp64_screened", G.PREC_FP64, args.variant_screen),
                             ("mixed", G.PREC_MIXED, 0.0),
                             ("mixed_screened", G.PREC_MIXED, args.variant_screen)):
            if name != "fp64" and args.no_mixed:
                continue
            en, res, _ = G.scf_run(mol, nrad, nang, lmax, threshold=1e-6, precision=pr, screen=sc)
            scf[name] = {"energy_Ha": res["energy"], "iterations": res["iterations"],
                         "converged": bool(res["converged"]), "last_dE": res["delta"],
                         "s_per_iter": res["seconds_iter"], "setup_s": res["seconds_setup"],
                         "device_step": bool(res["device_step"]), "host_fallbacks": res["host_fallbacks"]}
        for name in ("fp64_screened", "mixed", "mixed_screened"):
            # a difference of energies means something only when both runs converged
            # (the reference SCF oscillates without settling on the long alkanes)
            if name in scf:
                both = scf[name]["converged"] and scf["fp64"]["converged"]
                scf[name]["energy_error_vs_fp64_Ha"] = (scf[name]["energy_Ha"] - scf["fp64"]["energy_Ha"]) if both else None
        # the same iterations with the host eigen-solver step (s/iteration comparison)
        _, rh, _ = G.scf_run(mol, nrad, nang, lmax, max_iter=8, threshold=0.0, density_step=G.STEP_HOST)
        scf["fp64_host_step_8_iterations"] = {"s_per_iter": rh["seconds_iter"], "device_step": False}
        scf["grid_build_ms"] = ms_per_step
        out["scf"] = scf

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(mol, nrad, nang, lmax, P, units)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
