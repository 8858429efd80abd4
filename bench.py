#!/usr/bin/env python3
"""Benchmark: ns/day and CPIM solve ms/step of the MI355X force + CPIM path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1]
                    [--precision fp64|fp32|both] [--no-cpu]

Workload (default): Au561 Mackay icosahedron (6 shells) in a 6 nm box with
SPC/E water at 1.0 g/cm^3 (the seeded config-1 deck of capacmd_amd.gen),
Sutton-Chen gold, DSF water electrostatics (alpha 2 nm^-1, rc 1.2 nm), CPIM
on the gold (2,245 unknowns), Nose-Hoover NVT, dt = 1 fs.  A "step" is one
full MD step on the device: NHC, velocity Verlet + RATTLE, neighbour build,
pair/Sutton-Chen/CPIM forces with the BiCGSTAB solve, the molecule wrap at
the save cadence.  Each path runs with its default warm start: the previous
solution for FP64 (the reference's), quadratic extrapolation of the last
three solutions for FP32-refined (same stop test, fewer iterations); the
FP32-refined path is also timed with the previous solution as its start.  Inputs are resident in HBM before the timed region; the
library returns only after each step's work has finished (it synchronises on
the solver's convergence flag), so wall time around the K steps is device
time.  ns/day = 0.0864 / (seconds per step) at dt = 1 fs.

Prints one JSON line (rank 0; there is one GPU).  The roofline object is for
the relay SpMV, the kernel inside the solve (timed live with HIP events on
the library's stream); cpu_baseline times the reference program built from
its own sources (oracle/_ref, one MPI rank on one core) on a one-step sample
of the same deck, or the oracle restatement where that binary cannot run.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from capacmd_amd import gen, lib  # noqa: E402

HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(nm: int, n_mat: int, w: int) -> float:
    """Algorithmic bytes of one FP64 relay SpMV: the unique entries of the
    symmetric relay, w (8 N_m^2 - 3 N_m), plus 16 n_mat for the vectors."""
    return w * (8.0 * nm * nm - 3.0 * nm) + 16.0 * n_mat


def spmv32_bytes(nm: int, n_mat: int) -> float:
    """Bytes one FP32-refined SpMV must read: its relay format, 7 FP32
    coefficients per ordered metal pair (28 N_m^2 with the zero diagonal
    slots), plus 16 n_mat for the vectors.  This is below the unique-entry
    count of the FP32 matrix, 4 (8 N_m^2 - 3 N_m), because the format stores
    the block's generators rather than its 16 entries."""
    return 28.0 * nm * nm + 16.0 * n_mat


def traffic_from_profiles(kernel: str, algorithmic_bytes: float):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), as committed under
    profiles/; None when no such summary exists or when it was measured on
    another workload (its algorithmic bytes per launch differ)."""
    path = os.path.join(PROFILES, "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as fh:
            data = json.load(fh)
        entry = data.get(kernel, {})
        if int(entry.get("algorithmic_bytes", -1)) != int(algorithmic_bytes):
            return None
        return entry.get("bytes_per_launch")
    except Exception:
        return None


def run_gpu(deck, precision: int, steps: int, warmup: int, predictor=None, kernels=True):
    ctx = lib.Context(deck, precision=precision)
    if predictor is not None:
        ctx.set_predictor(predictor)
    ctx.force()
    ctx.md_begin()
    n_save = max(deck.n_save, 1)
    step = 1
    if warmup:
        done = ctx.md_run(step, warmup, n_save)
        step += done
    # the timed steps run as the product runs them (no per-phase timers, so
    # unobserved steps end without a host synchronisation)
    builds0 = ctx.nb_builds()
    t0 = time.perf_counter()
    done = ctx.md_run(step, steps, n_save)
    t1 = time.perf_counter()
    assert done == steps, done
    builds = ctx.nb_builds() - builds0
    step += done
    # the phase breakdown from as many further steps with the library's
    # HIP-event timers on (they synchronise at every step's end)
    ctx.reset_timing()
    ctx.enable_timing(True)
    done = ctx.md_run(step, steps, n_save)
    assert done == steps, done
    t, calls, iters = ctx.timing()
    obs = ctx.observables()
    ctx.enable_timing(False)
    sec = (t1 - t0) / steps
    out = {
        "ms_per_step": 1e3 * sec,
        "ns_per_day": 0.0864 / sec,
        "solve_ms_per_step": 1e3 * t[7] / steps,
        "matrix_ms_per_step": 1e3 * t[5] / steps,
        "pair_ms_per_step": 1e3 * t[4] / steps,
        "neighbour_ms_per_step": 1e3 * t[10] / steps,
        "cpim_force_ms_per_step": 1e3 * t[8] / steps,
        "sc_ms_per_step": 1e3 * t[2] / steps,
        "rhs_ms_per_step": 1e3 * t[6] / steps,
        "integrator_ms_per_step": 1e3 * t[12] / steps,
        "solve_iters_per_step": iters / steps,
        "nb_builds_per_step": builds / steps,
        "nb_skin_nm": ctx.skin,
        "warm_start_predictor": ctx.predictor,
        "final_temp": obs["temp"],
        "final_energy": float(obs["pot"][0]),
    }
    if not kernels:
        ctx.close()
        return out
    kern = {}
    kernels = [("spmv", lib.KERNEL_SPMV), ("assemble", lib.KERNEL_ASSEMBLE),
               ("pair", lib.KERNEL_PAIR), ("neighbour", lib.KERNEL_NEIGHBOUR),
               ("cpim_mm", lib.KERNEL_CPIM_MM)]
    if precision == lib.PREC_FP32_REFINED:
        kernels.append(("apply64", lib.KERNEL_APPLY64))
    kernels.append(("solve", lib.KERNEL_SOLVE))       # last: from the last MD step's warm start
    for name, kid in kernels:
        kern[name] = ctx.time_kernel(kid, 10 if kid == lib.KERNEL_SOLVE else 20)
    out["kernel_ms"] = kern
    ctx.close()
    return out


def fp32_error(deck, states: int = 3, stride: int = 20, crit: float = 1e-8):
    """Per-step error of the FP32-refined path against the FP64 path.

    The FP64 path runs MD; at `states` points `stride` steps apart, the
    FP32-refined path evaluates one force call from the identical state and
    the identical warm start (the FP64 solution of the step before).  Both
    solves stop at the reference's criterion (rTr < 1e-8 bTb), so part of
    the difference is the solver tolerance itself; the FP64 path agrees with
    the CPU oracle to 1e-10 at a given solution (tests/test_gpu_parity.py).
    Returns the largest relative errors (2-norms) of the induced charges,
    dipoles, all forces, metal forces and the potential energy."""
    c64 = lib.Context(deck, precision=lib.PREC_FP64)
    c32 = lib.Context(deck, precision=lib.PREC_FP32_REFINED)
    c64.set_solve_tolerance(crit)
    c32.set_solve_tolerance(crit)
    nm, n = deck.n_metal, deck.n_atoms
    worst = {"q": 0.0, "p": 0.0, "f": 0.0, "f_metal": 0.0, "epot": 0.0}

    def rel(a, b):
        den = float(np.linalg.norm(b))
        return float(np.linalg.norm(a - b)) / (den if den > 0 else 1.0)

    try:
        c64.force()
        c64.md_begin()
        step = 1
        for _ in range(states):
            if stride > 1:
                c64.md_run(step, stride - 1, 10 ** 9)
                step += stride - 1
            r_prev, v_prev, x_prev = c64.get_state()
            c64.md_run(step, 1, 10 ** 9)
            step += 1
            r, v, x64 = c64.get_state()
            f64 = c64.forces()
            e64 = float(np.sum(c64.observables()["pot"][1:]))
            c32.set_state(r=r, v=v, x=x_prev)
            f32, pot32, _ = c32.force()
            _, _, x32 = c32.get_state()
            e32 = float(np.sum(pot32[1:]))
            F64, F32 = f64.reshape(3, n), f32.reshape(3, n)
            errs = {"q": rel(x32[3 * nm:4 * nm], x64[3 * nm:4 * nm]),
                    "p": rel(x32[:3 * nm], x64[:3 * nm]),
                    "f": rel(f32, f64),
                    "f_metal": rel(F32[:, :nm], F64[:, :nm]),
                    "epot": abs(e32 - e64) / max(abs(e64), 1e-300)}
            for k, e in errs.items():
                worst[k] = max(worst[k], e)
    finally:
        c64.close()
        c32.close()
    return {k: float("%.3e" % v) for k, v in worst.items()} | {
        "states": states, "stride_steps": stride, "solve_crit": crit}


def cpu_reference(deck_dir: str, steps: int = 1):
    """The reference program on the host: one MPI rank, `steps` MD steps."""
    exe = os.path.join(ROOT, "oracle", "_ref", "capacmd_ref")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as tmp:
        for f in ("init.gro", "param.txt", "mdset.txt"):
            shutil.copy(os.path.join(deck_dir, f), tmp)
        mds = os.path.join(tmp, "mdset.txt")
        txt = open(mds).read()
        txt = re.sub(r"number_of_steps\s+\d+", f"number_of_steps   {steps}", txt)
        txt = re.sub(r"save_step\s+\d+", "save_step         1000000", txt)
        open(mds, "w").write(txt)
        env = dict(os.environ, MALLOC_MMAP_THRESHOLD_="0")
        try:
            res = subprocess.run([exe, "-gro", "init.gro", "-par", "param.txt", "-mds", "mdset.txt"],
                                 cwd=tmp, env=env, capture_output=True, text=True, timeout=300)
        except Exception as e:  # noqa: BLE001
            log("reference program did not run:", e)
            return None
        if res.returncode != 0:
            log("reference program failed:", res.stdout[-500:], res.stderr[-500:])
            return None
        m = re.search(r"Total time used in force:\s+([0-9.]+)", res.stdout)
        s = re.search(r"Time used in CPIM-solve:\s+([0-9.]+)", res.stdout)
        if not m:
            return None
        calls = steps + 1                       # the initial force call + one per step
        sec = float(m.group(1)) / calls
        return {"value": 0.0864 / sec, "unit": "ns/day", "cores": 1, "kind": "reference",
                "ms_per_step": 1e3 * sec,
                "solve_ms_per_step": 1e3 * float(s.group(1)) / calls if s else None,
                "sample": f"reference program (oracle/_ref/capacmd_ref, 1 MPI rank), same deck, "
                          f"{steps} MD step(s) = {calls} force calls incl. the first solve from zero"}


def cpu_port(deck, steps: int = 1):
    """The oracle restatement on one core, same sample."""
    from oracle.oracle import Oracle
    o = Oracle(deck)
    t0 = time.perf_counter()
    o.force()
    o.md_begin()
    for s in range(1, steps + 1):
        o.md_step(s, 10 ** 9)
    sec = (time.perf_counter() - t0) / (steps + 1)
    t = o.timers()
    return {"value": 0.0864 / sec, "unit": "ns/day", "cores": 1, "kind": "port",
            "ms_per_step": 1e3 * sec, "solve_ms_per_step": 1e3 * t[7] / (steps + 1),
            "sample": f"oracle restatement, 1 core, same deck, {steps} MD step(s) = "
                      f"{steps + 1} force calls"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "fp32", "both"], default="both")
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the whole box lives on one MI355X)", file=sys.stderr)
        return 2
    deck_spec = gen.config(a.config)
    tmp = tempfile.mkdtemp(prefix="capac_bench_")
    gen.write_deck(deck_spec, tmp)
    deck = lib.Deck.from_dir(tmp)
    log(f"deck {deck_spec.name}: N={deck.n_atoms} Nm={deck.n_metal} n_mat={deck.n_mat}")

    paths = {}
    precs = {"fp64": lib.PREC_FP64, "fp32": lib.PREC_FP32_REFINED}
    wanted = ["fp64", "fp32"] if a.precision == "both" else [a.precision]
    for name in wanted:
        try:
            paths[name] = run_gpu(deck, precs[name], a.steps, a.warmup)
            log(name, json.dumps(paths[name]))
        except lib.CapacError as e:
            log(f"{name} path unavailable: {e}")
    if not paths:
        print("no GPU path ran", file=sys.stderr)
        return 1
    if "fp32" in paths and paths["fp32"]["warm_start_predictor"] != 0:
        # the same path with the reference's warm start (previous solution),
        # for the predictor's share of the speed
        try:
            paths["fp32_previous_solution_start"] = run_gpu(deck, lib.PREC_FP32_REFINED, a.steps, a.warmup,
                                                            predictor=0, kernels=False)
            log("fp32 with the previous solution as start", json.dumps(paths["fp32_previous_solution_start"]))
        except lib.CapacError as e:
            log(f"fp32 run with predictor 0 failed: {e}")
    best = "fp32" if "fp32" in paths else "fp64"
    main_path = paths[best]

    nm, n_mat = deck.n_metal, deck.n_mat
    spmv_ms = main_path["kernel_ms"]["spmv"]
    nbytes = spmv_bytes(nm, n_mat, 8) if best == "fp64" else spmv32_bytes(nm, n_mat)
    achieved = nbytes / (spmv_ms * 1e-3) / 1e9
    roofline = {"kernel": "relay SpMV, " + ("k_spmv<RelayF64>" if best == "fp64" else "k_spmv<RelayC32>"),
                "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles("spmv_" + best, nbytes),
                "algorithmic_bytes_per_launch": nbytes,
                "ms_per_launch": spmv_ms}
    for name, p in paths.items():
        if "kernel_ms" not in p:
            continue
        b = spmv_bytes(nm, n_mat, 8) if name == "fp64" else spmv32_bytes(nm, n_mat)
        p["spmv_GBps"] = round(b / (p["kernel_ms"]["spmv"] * 1e-3) / 1e9, 1)
    speedup = None
    err32 = None
    if "fp32" in paths and "fp64" in paths:
        # the solve alone (the reference's timer category 7), from the last MD
        # step's system and warm start; in MD the pair forces overlap it
        speedup = round(paths["fp64"]["kernel_ms"]["solve"] / paths["fp32"]["kernel_ms"]["solve"], 3)
        # at the reference's stop test, and with both solves taken far below
        # it, so that what remains is the FP32 relay's own error
        err32 = {"note": "FP32-refined vs FP64 path from identical states and warm starts, max "
                         "over states; solve_crit is the stop test rTr < crit bTb",
                 "reference_tolerance": None, "tight_tolerance": None}
        for key, crit in (("reference_tolerance", 1e-8), ("tight_tolerance", 1e-11)):
            try:
                err32[key] = fp32_error(deck, crit=crit)
            except lib.CapacError as e:
                log(f"fp32 error measurement at crit {crit} failed: {e}")
        log("fp32-refined error vs fp64:", json.dumps(err32))

    cpu = None
    if not a.no_cpu and deck.n_atoms <= 46341:
        cpu = cpu_reference(tmp, 1)
        if cpu is None:
            cpu = cpu_port(deck, 1)
    shutil.rmtree(tmp, ignore_errors=True)

    line = {
        "metric": "ns/day + CPIM solve ms/step (Au NP in water); FP32-refined vs FP64, 1 GPU",
        "value": round(main_path["ns_per_day"], 4),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(main_path["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best == "fp64" else "f32-matrix/f64-refined",
        "data": "synthetic (seeded generator: Mackay icosahedron + SPC/E lattice water)",
        "config": {"workload": f"{deck_spec.name}: Au{nm} + {deck_spec.n_water} SPC/E waters, "
                               f"{deck_spec.box[0]:g} nm box, {n_mat} CPIM unknowns, MD step dt 1 fs",
                   "atoms": deck.n_atoms, "path": best},
        "solve_ms_per_step": round(main_path["solve_ms_per_step"], 4),
        "solve_ms_alone": round(main_path["kernel_ms"]["solve"], 4),
        "solve_speedup_fp32_refined_vs_fp64": speedup,
        "fp32_refined_error": err32,
        "paths": paths,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
