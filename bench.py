#!/usr/bin/env python3
"""bench.py -- direct J+K builds/s and RHF gradient seconds on one MI355X,
FP64 and mixed precision.

Workload: (H2O)20 water cluster (seeded, 3.0 A lattice, +-0.3 A jitter, Haar
rotations), cc-pVDZ spherical (480 basis functions, 240 shells).

1. Two RHF runs through the C++ SCF driver (molecular::SCF::rhf over the HIP
   engine), each timed after an untimed warm-up run of both: plain FP64, then
   mixed precision (FP32 for quartets with
   Q_ab Q_cd max|D| below the mixed threshold and for the incremental
   Delta-D builds, FP64 for the last iterations).  Each is followed by the
   nuclear gradient (Fock::compute_forces).  Reports seconds, iterations,
   |dE| and max|dg| between the two.
2. The timed step: one full J+K build (Fock::formJKdirect semantics,
   src/fockbuild.cpp:64-164, precision 1e-12) at the converged FP64 density,
   with D, J and K resident in HBM.  `value` is FP64 builds/s; the mixed
   build rate and its error are reported beside it: FP32 below the split
   Q_ab Q_cd max|D| < t, reported at the SCF's own split t = 1e-4 (fixed
   before timing); the sweep t = 1e-4 .. 1e-1 is listed beside it.
   The nuclear gradient of the mixed SCF takes its quartets below 1e-4 in
   FP32 as well.
3. Roofline of the dominant class kernel of the FP64 build: algorithmic
   FLOPs from the committed op-count model (molecular2_amd/flop_model.json)
   over the kernel's average launch time, measured with HIP events recorded
   on the engine's stream around each launch.  HBM traffic per launch comes
   from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary when present.
4. cpu_baseline: the oracle's OpenMP FP64 restatement of formJKdirect timed on
   a bounded sample of the same build.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "J+K builds/s and RHF grad s: mixed vs FP64 GPU at |ΔE|<1µEh"
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector (= FP64 matrix) peak
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "jk_traffic_h2o20.json")


def class_flops(model, prof, pure=True):
    """FLOPs per class combo from the committed op-count model: `flops` over
    the primitive quartets the kernels compute (after primitive screening),
    `flops_all` over every primitive quartet of the surviving shell quartets
    (the unscreened textbook count)."""
    out = {}
    for (X, Y), (ms, nl, q, pq, cpq) in prof.items():
        m = model["classes"]["%d,%d" % (X, Y)]
        n_eri = m["n_eri_pure"] if pure else m["n_eri_cart"]
        fc = m["f_cont_pure"] if pure else m["f_cont_cart"]
        per_q = q * (fc + model["digest_flops_per_eri"] * n_eri)
        out[(X, Y)] = dict(ms=ms, launches=nl, quartets=q, prim_quartets=pq, computed_prim_quartets=cpq,
                           flops=cpq * m["f_prim"] + per_q, flops_all=pq * m["f_prim"] + per_q, cls=m["cls"],
                           eris=q * n_eri)
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mol, D, total_primq, budget_s=15.0):
    """Oracle FP64 restatement of formJKdirect (own OS/HGP ERIs, every
    primitive quartet) on bounded samples of the same build: the outer-shell
    indices s1 from the top down until the time budget is spent, once on all
    host threads (OpenMP over s1) and once on one thread (the reference's
    Fock build is serial).  A full build takes minutes on the CPU, so the
    rate is EXTRAPOLATED to builds/s by primitive-quartet count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    ob = O.Basis(mol.shells("cc-pvdz"), pure=True)
    Q = ob.schwarz()
    nsh = ob.nsh

    def sample(nthreads, budget):
        hi, done_q, done_pq, t_used = nsh, 0, 0, 0.0
        while hi > 0 and t_used < budget:
            lo = max(0, hi - (2 if nthreads > 1 else 1))
            t = time.perf_counter()
            q, pq = ob.jk_direct_range(D, Q, lo, hi, 1e-12, nthreads)
            t_used += time.perf_counter() - t
            done_q += q
            done_pq += pq
            hi = lo
        return hi, done_q, done_pq, t_used

    nthreads = min(16, os.cpu_count() or 1)
    hi, q, pq, t = sample(nthreads, budget_s)
    hi1, q1, pq1, t1 = sample(1, budget_s / 2)
    rate, rate1 = pq / t if t else 0.0, pq1 / t1 if t1 else 0.0
    return {"value": rate / total_primq if total_primq else None, "unit": "J+K builds/s",
            "cores": nthreads, "kind": "port",
            "value_1thread": rate1 / total_primq if total_primq else None,
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
            "sample": "oracle formJKdirect restatement (own OS/HGP ERIs, not libint2) on outer shells "
                      "s1 in [%d, %d) of the same build with %d threads (%d quartets / %d primitive quartets in "
                      "%.1f s) and s1 in [%d, %d) on 1 thread (%d / %d in %.1f s); builds/s EXTRAPOLATED by "
                      "primitive-quartet count (the build has %d)" %
                      (hi, nsh, nthreads, q, pq, t, hi1, nsh, q1, pq1, t1, total_primq)}


def df_bench(eng, mol, Cocc, n, steps, torch, dev):
    """Density-fitted J/K (compute_2body_fock_df, src/fockbuild.cpp:585-660)
    with cc-pVDZ-JKFIT on the converged occupied orbitals: fitted-tensor setup
    time, and per call the FP32 (MFMA) and FP64 GEMM paths.  FLOPs per call
    4 naux n^2 nocc (X = C^T B and K = X^T X) + 4 naux n(n+1)/2 (J)."""
    aux = mol.shells("cc-pvdz-jkfit")
    eng.set_basis(aux, pure=True, which=1)
    naux = sum(2 * s.l + 1 for s in aux)
    nocc = Cocc.shape[1]
    torch.cuda.synchronize()
    t = time.perf_counter()
    eng.df_setup()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t
    dC = torch.from_numpy(np.ascontiguousarray(Cocc.T)).to(dev)   # column-major n x nocc
    dJ = torch.empty((n, n), dtype=torch.float64, device=dev)
    dK = torch.empty_like(dJ)
    flops = 4.0 * naux * n * n * nocc + 4.0 * naux * n * (n + 1) / 2
    res = {"naux": naux, "setup_s": round(t_setup, 3), "flops_per_call": flops}
    for name, fp32, peak in (("fp32", True, 157.3), ("fp64", False, 78.6)):
        eng.jk_df_device(dC.data_ptr(), nocc, dJ.data_ptr(), dK.data_ptr(), fp32)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            eng.jk_df_device(dC.data_ptr(), nocc, dJ.data_ptr(), dK.data_ptr(), fp32)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t) / steps
        eng.set_profiling(True)
        eng.jk_df_device(dC.data_ptr(), nocc, dJ.data_ptr(), dK.data_ptr(), fp32)
        eng.set_profiling(False)
        tx, tk, tj, tt = eng.df_timing()
        res[name] = {"ms_per_call": round(ms, 3), "gemm_X_ms": round(tx, 3), "gemm_K_ms": round(tk, 3),
                     "J_ms": round(tj, 3), "tflops": round(flops / (ms * 1e-3) / 1e12, 2),
                     "matrix_peak_tflops": peak}
        res[name]["J"], res[name]["K"] = dJ.cpu().numpy(), dK.cpu().numpy()
    J32, K32, J64, K64 = res["fp32"].pop("J"), res["fp32"].pop("K"), res["fp64"].pop("J"), res["fp64"].pop("K")
    P = Cocc @ Cocc.T
    # two-electron energy of the fitted build at these orbitals, tr(P (2J - K))
    e64, e32 = float(np.sum(P * (2 * J64 - K64))), float(np.sum(P * (2 * J32 - K32)))
    res["fp32_vs_fp64"] = {"speedup": round(res["fp64"]["ms_per_call"] / res["fp32"]["ms_per_call"], 3),
                           "max_abs_dJ": float(np.max(np.abs(J32 - J64))),
                           "max_abs_dK": float(np.max(np.abs(K32 - K64))),
                           "E2_Eh": e64, "abs_dE2_Eh": abs(e32 - e64),
                           "note": "FP32 fits at fixed orbitals; the mixed DF SCF finishes with FP64 iterations"}
    return res


def batch_bench(eng, nmol=8):
    """Batched SCF (m2_rhf_batch_min) on the first `nmol` molecules of the
    seeded 256-molecule H/C/N/O batch (3-30 atoms, cc-pVDZ), FP64 and mixed,
    from the SOAD guess (the reference's default); molecules/s counts every
    molecule of the sample, converged or not.  maxiter is raised to 100 (the
    random clusters need 50-100 iterations); the other options are the
    reference's defaults.  A core-Hamiltonian-guess FP64 run is listed beside
    them."""
    from molecular2_amd import molecules as M
    mols = M.random_batch(256, seed=256)[:nmol]
    r64, w64 = eng.rhf_batch(mols, nworkers=8, maxiter=100, guess="soad")
    rmx, wmx = eng.rhf_batch(mols, nworkers=8, maxiter=100, mixed=1, guess="soad")
    rhc, whc = eng.rhf_batch(mols, nworkers=8, maxiter=100, guess="hcore")
    both = [(a["energy"], b["energy"]) for a, b in zip(r64, rmx) if a["converged"] and b["converged"]]
    dE = max((abs(x - y) for x, y in both), default=None)
    return {"sample": "first %d of random_batch(256, seed=256): %s atoms" % (nmol, [m.natom for m in mols]),
            "workers": 8,
            "fp64": {"molecules_per_s": round(nmol / w64, 3), "wall_s": round(w64, 2),
                     "converged": sum(r["converged"] for r in r64), "iterations": sum(r["iterations"] for r in r64)},
            "mixed": {"molecules_per_s": round(nmol / wmx, 3), "wall_s": round(wmx, 2),
                      "converged": sum(r["converged"] for r in rmx), "iterations": sum(r["iterations"] for r in rmx)},
            "fp64_core_guess": {"molecules_per_s": round(nmol / whc, 3), "wall_s": round(whc, 2),
                                "converged": sum(r["converged"] for r in rhc),
                                "iterations": sum(r["iterations"] for r in rhc)},
            "max_abs_dE_Eh_converged": dE}


def traffic_for(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k.get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nwater", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-scf", action="store_true",
                    help="skip the SCF/gradient runs (J/K timing at the core-guess density)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-SCF sample")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 256 MiB device copy (HBM byte-counter calibration for rocprofv3 --pmc)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(the J/K quartet list would be split across GPUs with one J/K all-reduce).", file=sys.stderr)
        sys.exit(2)

    import torch
    from molecular2_amd import molecules as M
    from molecular2_amd.engine import Engine

    dev = torch.device("cuda:0")
    if a.copy_calibration:
        src = torch.ones(32 << 20, dtype=torch.float64, device=dev)   # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    mol = M.water_cluster(a.nwater, seed=20)
    shells = mol.shells("cc-pvdz")
    eng = Engine(0)
    out_scf = None
    if not a.no_scf:
        # SOAD start: the reference's default guess (src/ProgramController.cpp:307)
        kw = dict(enconverge=1e-8, densconverge=1e-5, maxiter=60, guess="soad")
        # untimed warm-up of both: the first launches load the code objects and
        # create the library handles
        eng.rhf(mol, "cc-pvdz", pure=True, gradient=True, **kw)
        eng.rhf(mol, "cc-pvdz", pure=True, gradient=True, mixed=1, **kw)
        r64 = eng.rhf(mol, "cc-pvdz", pure=True, gradient=True, **kw)
        rmx = eng.rhf(mol, "cc-pvdz", pure=True, gradient=True, mixed=1, **kw)
        D = r64["D"]
        out_scf = {
            "fp64": {"energy_Eh": r64["energy"], "iterations": r64["iterations"], "converged": r64["converged"],
                     "scf_s": round(r64["total_seconds"] - r64["grad_seconds"], 3),
                     "jk_s": round(r64["jk_seconds"], 3), "grad_s": round(r64["grad_seconds"], 3)},
            "mixed": {"energy_Eh": rmx["energy"], "iterations": rmx["iterations"], "converged": rmx["converged"],
                      "fp32_builds": rmx["n_fp32"],
                      "scf_s": round(rmx["total_seconds"] - rmx["grad_seconds"], 3),
                      "jk_s": round(rmx["jk_seconds"], 3), "grad_s": round(rmx["grad_seconds"], 3)},
            "abs_dE_Eh": abs(rmx["energy"] - r64["energy"]),
            "max_abs_dgrad_Eh_per_bohr": float(np.max(np.abs(rmx["grad"] - r64["grad"]))),
            "scf_speedup": round((r64["total_seconds"] - r64["grad_seconds"]) /
                                 max(1e-9, rmx["total_seconds"] - rmx["grad_seconds"]), 3),
            "grad_speedup": round(r64["grad_seconds"] / max(1e-9, rmx["grad_seconds"]), 3),
            "note": "the mixed gradient takes the derivative quartets whose bound is below 1e-4 in FP32, "
                    "each gradient on its own SCF's density",
        }
    t = time.perf_counter()
    eng.set_basis(shells, pure=True)
    t_setup = time.perf_counter() - t
    if a.no_scf:
        S, T, V = eng.one_body(mol.Z, mol.xyz)
        w, U = np.linalg.eigh(S)
        X = U @ np.diag(w ** -0.5) @ U.T
        _, C = np.linalg.eigh(X.T @ (T + V) @ X)
        Co = (X @ C)[:, :mol.nel // 2]
        D = 2.0 * Co @ Co.T
    n = eng.nbf
    dD = torch.from_numpy(np.ascontiguousarray(D)).to(dev)
    dJ = torch.empty_like(dD)
    dK = torch.empty_like(dD)
    torch.cuda.synchronize()

    def run(steps, mixed):
        for _ in range(steps):
            eng.jk_device(dD.data_ptr(), dJ.data_ptr(), dK.data_ptr(), 1.0, 1e-12, mixed)

    # FP64 path (timed region: K builds, HIP events around each class launch)
    run(a.warmup, False)
    eng.reset_profile()
    eng.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(a.steps, False)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    eng.set_profiling(False)
    prof_t = eng.profile()
    J64 = dJ.cpu().numpy()
    K64 = dK.cpu().numpy()
    # computed quartets / primitive quartets per class: device counters on one
    # extra (untimed) build of the same density
    eng.reset_profile()
    eng.set_profiling(True, counters=True)
    run(1, False)
    eng.set_profiling(False)
    prof_c = eng.profile()
    prof = {k: (v[0], v[1], prof_c[k][2] * a.steps, prof_c[k][3] * a.steps, prof_c[k][4] * a.steps)
            for k, v in prof_t.items()}
    # mixed path: FP32 below the split Q_ab Q_cd max|D| < t.  Each split of
    # the sweep is timed; the reported mixed build is the fastest whose
    # two-electron energy stays within 1 uEh of the FP64 build's (the
    # library's SCF default split is 1e-4)
    e64 = 0.5 * np.sum(D * (J64 - 0.5 * K64))
    sweep = []
    for thr in (1e-4, 1e-3, 1e-2, 1e-1):
        eng.set_mixed_threshold(thr)
        run(a.warmup, True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(a.steps, True)
        torch.cuda.synchronize()
        tt = time.perf_counter() - t0
        Jt, Kt = dJ.cpu().numpy(), dK.cpu().numpy()
        de = float(abs(0.5 * np.sum(D * (Jt - 0.5 * Kt)) - e64))
        sweep.append({"fp32_below": thr, "ms": round(1e3 * tt / a.steps, 3), "abs_dE2_Eh": de,
                      "max_abs_dJ": float(np.max(np.abs(Jt - J64))), "max_abs_dK": float(np.max(np.abs(Kt - K64))),
                      "_t": tt, "_J": Jt, "_K": Kt})
    # per-class FP32 / FP64 kernel time: one profiled build with every
    # quartet in FP32 (timing only: the FP32 split of the mixed build keeps
    # the large quartets in FP64)
    eng.set_mixed_threshold(1e300)
    run(a.warmup, True)
    eng.reset_profile()
    eng.set_profiling(True)
    run(a.steps, True)
    torch.cuda.synchronize()
    eng.set_profiling(False)
    prof32 = eng.profile()
    # the reported mixed build uses the SCF's own split, fixed before timing
    fixed = next(x for x in sweep if x["fp32_below"] == 1e-4)
    tmx, Jm, Km, thr_mx = fixed["_t"], fixed["_J"], fixed["_K"], fixed["fp32_below"]
    best = min((x for x in sweep if x["abs_dE2_Eh"] < 1e-6), key=lambda x: x["_t"])
    eng.set_mixed_threshold(1e-4)
    sweep = [{k: v for k, v in x.items() if not k.startswith("_")} for x in sweep]

    with open(os.path.join(ROOT, "molecular2_amd", "flop_model.json")) as fh:
        model = json.load(fh)
    cf = class_flops(model, prof)
    tot_flops = sum(v["flops"] for v in cf.values()) / a.steps
    tot_primq = sum(v["prim_quartets"] for v in cf.values()) / a.steps
    tot_cprimq = sum(v["computed_prim_quartets"] for v in cf.values()) / a.steps
    tot_flops_all = sum(v["flops_all"] for v in cf.values()) / a.steps
    tot_q = sum(v["quartets"] for v in cf.values()) / a.steps
    tot_eri = sum(v["eris"] for v in cf.values()) / a.steps
    dom = max(cf, key=lambda k: cf[k]["ms"])
    d = cf[dom]
    ms_launch = d["ms"] / max(1, d["launches"])
    achieved = d["flops"] / max(1, d["launches"]) / (ms_launch * 1e-3) / 1e12
    kname = "k_jk%s double" % d["cls"]

    builds64 = a.steps / t64
    buildsmx = a.steps / tmx
    # two-electron energy of the build at this density: E2 = 1/2 tr(D (J - K/2))
    emx = 0.5 * np.sum(D * (Jm - 0.5 * Km))
    out = {
        "metric": METRIC,
        "value": round(builds64, 4),
        "unit": "J+K builds/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * t64 / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "(H2O)%d cc-pVDZ direct J+K build (FP64), %s density" %
                               (a.nwater, "core-guess" if a.no_scf else "converged RHF"),
                   "molecule": mol.name, "nbf": n, "nshells": len(shells), "precision": 1e-12},
        "roofline": {"bound": "valu", "kernel": kname,
                     "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": traffic_for(kname),
                     "ms_per_launch": round(ms_launch, 4),
                     "flops_per_launch": d["flops"] / max(1, d["launches"]),
                     "flops_basis": "op-count model over the primitive quartets computed (after primitive "
                                    "screening)",
                     "achieved_all_prim_quartets": round(d["flops_all"] / max(1, d["launches"]) /
                                                         (ms_launch * 1e-3) / 1e12, 3)},
        "fp64": {"builds_per_s": round(builds64, 4), "ms": round(1e3 * t64 / a.steps, 3),
                 "algorithmic_tflops": round(tot_flops / (t64 / a.steps) / 1e12, 3),
                 "quartets": int(tot_q), "prim_quartets": int(tot_primq),
                 "computed_prim_quartets": int(tot_cprimq),
                 "tflops_all_prim_quartets": round(tot_flops_all / (t64 / a.steps) / 1e12, 3),
                 "eris": int(tot_eri),
                 "eris_per_s": round(tot_eri / (t64 / a.steps), 1)},
        "mixed": {"builds_per_s": round(buildsmx, 4), "ms": round(1e3 * tmx / a.steps, 3),
                  "speedup_vs_fp64": round(buildsmx / builds64, 3), "fp32_below": thr_mx,
                  "max_abs_dJ": float(np.max(np.abs(Jm - J64))), "max_abs_dK": float(np.max(np.abs(Km - K64))),
                  "abs_dE2_Eh": float(abs(emx - e64)), "sweep": sweep,
                  "fastest_split_within_1uEh": best["fp32_below"],
                  "fastest_speedup_within_1uEh": round(t64 / best["_t"], 3),
                  "class_fp64_over_fp32_time": {
                      cf[k]["cls"]: round(cf[k]["ms"] / max(prof32[k][0], 1e-9), 3) for k in cf if k in prof32},
                  "class_ratio_note": "per class combination, FP64 kernel time over its time with every quartet in "
                                      "FP32 (a timing-only build)"},
        "classes": {v["cls"]: {"ms_per_build": round(v["ms"] / a.steps, 3),
                               "tflops": round(v["flops"] / max(v["ms"], 1e-9) / 1e9, 3),
                               "quartets": v["quartets"] // a.steps} for v in cf.values()},
        "setup_s": round(t_setup, 3),
    }
    if out_scf is not None:
        out["scf"] = out_scf
        out["df"] = df_bench(eng, mol, r64["C"][:, :mol.nel // 2], n, a.steps, torch, dev)
    if not a.no_batch and not a.no_scf:
        out["batch"] = batch_bench(eng)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mol, D, tot_primq, a.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
