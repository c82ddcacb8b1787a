#!/usr/bin/env python3
"""bench.py -- direct J+K builds/s on one MI355X.

Workload: (H2O)20 water cluster (seeded, 3.0 A lattice, +-0.3 A jitter, Haar
rotations), cc-pVDZ spherical (480 basis functions, 240 shells), Schwarz x
density screening at precision 1e-12, one full J+K build per step
(Fock::formJKdirect semantics, src/fockbuild.cpp:64-164).  The density is
the idempotent core-Hamiltonian-guess density 2 C_occ C_occ^T of the same
molecule (no converged density is needed to exercise every quartet class).
Inputs are resident in HBM before timing starts.

Prints ONE JSON line.  Also reports the FP64 and mixed-precision build rates,
the roofline fraction of the dominant class kernel (FP64 VALU peak) and the
oracle's CPU FP64 restatement timed on a bounded sample of the same build.
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
FP32_PEAK_TFLOPS = 157.3


def core_guess_density(S, H, nocc):
    w, U = np.linalg.eigh(S)
    X = U @ np.diag(w ** -0.5) @ U.T
    e, C = np.linalg.eigh(X.T @ H @ X)
    C = X @ C
    Co = C[:, :nocc]
    return 2.0 * Co @ Co.T


def class_flops(model, prof, pure=True):
    """algorithmic FLOPs per class combo from the committed op-count model."""
    out = {}
    for (X, Y), (ms, nl, q, pq) in prof.items():
        m = model["classes"]["%d,%d" % (X, Y)]
        n_eri = m["n_eri_pure"] if pure else m["n_eri_cart"]
        fc = m["f_cont_pure"] if pure else m["f_cont_cart"]
        flops = pq * m["f_prim"] + q * (fc + model["digest_flops_per_eri"] * n_eri)
        out[(X, Y)] = dict(ms=ms, launches=nl, quartets=q, prim_quartets=pq, flops=flops, cls=m["cls"],
                           eris=q * n_eri)
    return out


def cpu_baseline(mol, D, total_primq, budget_s=15.0):
    """Oracle FP64 restatement of formJKdirect on a bounded sample (the
    heaviest outer-shell indices s1 of the same build), extrapolated by
    primitive-quartet count to builds/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    ob = O.Basis(mol.shells("cc-pvdz"), pure=True)
    Q = ob.schwarz()
    nthreads = min(16, os.cpu_count() or 1)
    nsh = ob.nsh
    hi = nsh
    done_q = done_pq = 0
    t_used = 0.0
    while hi > 0 and t_used < budget_s:
        lo = max(0, hi - 2)
        t = time.perf_counter()
        q, pq = ob.jk_direct_range(D, Q, lo, hi, 1e-12, nthreads)
        t_used += time.perf_counter() - t
        done_q += q
        done_pq += pq
        hi = lo
    rate = done_pq / t_used if t_used > 0 else 0.0
    return {"value": rate / total_primq if total_primq else None, "unit": "J+K builds/s",
            "cores": nthreads, "kind": "port",
            "sample": "oracle formJKdirect restatement (own OS/HGP ERIs, not libint2) over outer shells "
                      "s1 in [%d, %d) of the same build: %d quartets / %d primitive quartets in %.1f s, "
                      "extrapolated by primitive-quartet count" % (hi, nsh, done_q, done_pq, t_used)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nwater", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(the J/K quartet list would be split across GPUs with one J/K all-reduce).", file=sys.stderr)
        sys.exit(2)

    import torch
    from molecular2_amd import molecules as M
    from molecular2_amd.engine import Engine

    mol = M.water_cluster(a.nwater, seed=20)
    shells = mol.shells("cc-pvdz")
    eng = Engine(0)
    t = time.perf_counter()
    eng.set_basis(shells, pure=True)
    t_setup = time.perf_counter() - t
    S, T, V = eng.one_body(mol.Z, mol.xyz)
    D = core_guess_density(S, T + V, mol.nel // 2)
    n = eng.nbf
    dev = torch.device("cuda:0")
    dD = torch.from_numpy(D).to(dev)
    dJ = torch.empty_like(dD)
    dK = torch.empty_like(dD)
    torch.cuda.synchronize()

    def run(steps, mixed):
        for _ in range(steps):
            eng.jk_device(dD.data_ptr(), dJ.data_ptr(), dK.data_ptr(), 1.0, 1e-12, mixed)

    # FP64 path
    run(a.warmup, False)
    eng.reset_profile()
    eng.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(a.steps, False)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    eng.set_profiling(False)
    prof = eng.profile()
    J64 = dJ.cpu().numpy()
    K64 = dK.cpu().numpy()
    # mixed path
    run(a.warmup, True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(a.steps, True)
    torch.cuda.synchronize()
    tmx = time.perf_counter() - t0
    Jm = dJ.cpu().numpy()
    Km = dK.cpu().numpy()

    with open(os.path.join(ROOT, "molecular2_amd", "flop_model.json")) as fh:
        model = json.load(fh)
    cf = class_flops(model, prof)
    tot_flops = sum(v["flops"] for v in cf.values()) / a.steps
    tot_primq = sum(v["prim_quartets"] for v in cf.values()) / a.steps
    tot_q = sum(v["quartets"] for v in cf.values()) / a.steps
    tot_eri = sum(v["eris"] for v in cf.values()) / a.steps
    dom = max(cf, key=lambda k: cf[k]["ms"])
    d = cf[dom]
    ms_launch = d["ms"] / max(1, d["launches"])
    achieved = d["flops"] / max(1, d["launches"]) / (ms_launch * 1e-3) / 1e12

    builds64 = a.steps / t64
    buildsmx = a.steps / tmx
    # energy-type error of the mixed build at this density: E2 = 1/2 tr(D (J - K/2))
    e64 = 0.5 * np.sum(D * (J64 - 0.5 * K64))
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
        "config": {"workload": "(H2O)%d cc-pVDZ direct J+K build (FP64), core-guess density" % a.nwater,
                   "molecule": mol.name, "nbf": n, "nshells": len(shells), "precision": 1e-12},
        "roofline": {"bound": "valu", "kernel": "k_jk %s (FP64)" % d["cls"],
                     "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None,
                     "ms_per_launch": round(ms_launch, 4)},
        "fp64": {"builds_per_s": round(builds64, 4), "ms": round(1e3 * t64 / a.steps, 3),
                 "algorithmic_tflops": round(tot_flops / (t64 / a.steps) / 1e12, 3),
                 "quartets": int(tot_q), "prim_quartets": int(tot_primq), "eris": int(tot_eri),
                 "eris_per_s": round(tot_eri / (t64 / a.steps), 1)},
        "mixed": {"builds_per_s": round(buildsmx, 4), "ms": round(1e3 * tmx / a.steps, 3),
                  "speedup_vs_fp64": round(builds64 and buildsmx / builds64, 3),
                  "max_abs_dJ": float(np.max(np.abs(Jm - J64))), "max_abs_dK": float(np.max(np.abs(Km - K64))),
                  "abs_dE2_Eh": float(abs(emx - e64))},
        "classes": {v["cls"]: {"ms_per_build": round(v["ms"] / a.steps, 3),
                               "tflops": round(v["flops"] / max(v["ms"], 1e-9) / 1e9, 3),
                               "quartets": v["quartets"] // a.steps} for v in cf.values()},
        "setup_s": round(t_setup, 3),
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mol, D, tot_primq, a.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
