#!/usr/bin/env python3
"""Benchmark: ns/day and CPIM solve ms/step of the MI355X force + CPIM path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 3]
                    [--equil E] [--precision fp64|fp32|both] [--no-cpu] [--no-error]

Workload (default): Au2057 Mackay icosahedron (9 shells) in a 10 nm box with
SPC/E water at 1.0 g/cm^3 (the seeded config-3 deck of capacmd_amd.gen:
98,189 atoms, 8,229 CPIM unknowns), Sutton-Chen gold, DSF water
electrostatics (alpha 2 nm^-1, rc 1.2 nm), CPIM on the gold, Nose-Hoover NVT
at 300 K, dt = 1 fs.  A "step" is one full MD step on the device: NHC,
velocity Verlet + RATTLE, neighbour build, pair/Sutton-Chen/CPIM forces with
the CPIM solve (GMRES on the FP32 relay with FP64 refinement; the reference's
BiCGSTAB on the FP64 path), the molecule wrap at the save cadence.

Before anything is timed the deck is equilibrated for E MD steps
(FP32-refined path) so that the timed window runs near the thermostat's
300 K; both precisions then start from that state, thermostat included.
Each path runs with its default warm start: the previous solution for FP64
(the reference's), quadratic extrapolation of the last three solutions for
FP32-refined (same stop test, fewer iterations); the FP32-refined path is
also timed with the previous solution as its start.  Inputs are resident in
HBM before the timed region; the library returns only after the last step's
work has finished, so wall time around the K steps is device time.
ns/day = 0.0864 / (seconds per step) at dt = 1 fs.

Prints one JSON line (there is one GPU).  The roofline object is for the
relay SpMV, the HBM stream of the solve, timed live with HIP events on the
library's stream.  pair_valu gives the pair launches, which take the largest
share of the MD step's kernel time and are bound by FP64 VALU issue: row
entries and VALU instructions per second against the FP64 issue peak.
cpu_baseline
times the oracle restatement (OpenMP over the reference's decomposition
units, every core the process may use) on two MD steps from the same
equilibrated state and warm start.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
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
    """Algorithmic bytes of one relay SpMV: the unique entries of the
    symmetric relay, w (8 N_m^2 - 3 N_m), plus 16 n_mat for the vectors
    (w = 8 FP64, 4 FP32)."""
    return w * (8.0 * nm * nm - 3.0 * nm) + 16.0 * n_mat


def format_bytes(path: str, layout, nm: int, n_mat: int):
    """Bytes the SpMV's storage format moves per launch, plus 16 n_mat for
    the vectors: the FP32 relay's four coefficients (16 B) per ordered metal
    pair over rows padded to a multiple of 64 (ROWS), or per slot of the
    Nb (Nb + 1) / 2 tiles of 64 x 64 (TILES, Nb = ceil(N_m / 64)) plus its
    partial sums written and read once (2 x 2 KB per block pair); None for
    the matrix-free FP64 product (it reads positions and vectors only)."""
    if path != "fp32":
        return None
    if layout == "tiles":
        nb = (nm + 63) // 64
        return 16.0 * 4096 * nb * (nb + 1) / 2 + 4096.0 * nb * nb + 16.0 * n_mat
    ld = (nm + 63) // 64 * 64
    return 16.0 * nm * ld + 16.0 * n_mat


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


FP64_ISSUE_PEAK_G = 256 * 4 * 2.4 / 4        # wavefront FP64 VALU instructions per ns: 614.4 G/s


def pair_valu_roofline(path, deck):
    """The pair launches' FP64-VALU figures (they are VALU-bound, not
    HBM-bound): row entries per second and VALU instructions per second
    against the FP64 issue peak (one wavefront FP64 instruction per SIMD every
    4 cycles), the instruction count per launch from the committed rocprofv3
    summary of the same workload (None for another workload), and the HBM
    rate on the algorithmic bytes 4 P_list + 60 N."""
    ms = path.get("kernel_ms", {}).get("pair")
    entries = path.get("pair_row_entries")
    if not ms or not entries:
        return None
    out = {"kernels": "one force call's pair launches, alone on the full grid (k_pack_xyzq and "
                      "k_pair_water<true> on the Au/water decks; k_pair_rest too where a pair type is neither "
                      "funct 1 nor 4)",
           "bound": "fp64-valu", "ms_per_launch": ms, "row_entries": entries,
           "G_entries_per_s": round(entries / (ms * 1e-3) / 1e9, 1),
           "algorithmic_bytes": 4 * entries + 60 * deck.n_atoms,
           "hbm_GBps_algorithmic": round((4 * entries + 60 * deck.n_atoms) / (ms * 1e-3) / 1e9, 1),
           "peak_G_instr_per_s": FP64_ISSUE_PEAK_G}
    try:
        with open(os.path.join(PROFILES, "valu.json")) as fh:
            v = json.load(fh)
        if int(v.get("atoms", -1)) == deck.n_atoms and "k_pair_water<true>" in v:
            instr = v["k_pair_water<true>"]["valu_instr"] + v.get("k_pair_rest", {}).get("valu_instr", 0.0)
            out["valu_instr_per_launch"] = instr
            out["achieved_G_instr_per_s"] = round(instr / (ms * 1e-3) / 1e9, 1)
            out["frac"] = round(out["achieved_G_instr_per_s"] / FP64_ISSUE_PEAK_G, 4)
    except (OSError, ValueError, KeyError):
        pass
    return out


def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable": usable, "model": model}


def equilibrate(deck, steps: int):
    """MD on the FP32-refined path from the deck's start; returns the state
    (r, v, vec_pq), the thermostat chain and the temperature at the end."""
    ctx = lib.Context(deck, precision=lib.PREC_FP32_REFINED)
    ctx.force()
    ctx.md_begin()
    n_save = max(deck.n_save, 1)
    t0 = time.perf_counter()
    if steps:
        done = ctx.md_run(1, steps, n_save)
        assert done == steps, done
    t1 = time.perf_counter()
    r, v, x = ctx.get_state()
    obs = ctx.observables()
    ctx.close()
    log(f"equilibration: {steps} steps in {t1 - t0:.1f} s, T = {obs['temp']:.1f} K")
    return {"r": r, "v": v, "x": x, "vq": obs["vq"], "temp": obs["temp"], "steps": steps}


FP32_LAYOUT = lib.LAYOUT_AUTO        # --fp32-layout
BLOCK_PRECOND = None                  # --block-precond (None: the library's default)


def start_context(deck, precision: int, eq, predictor=None):
    ctx = lib.Context(deck, precision=precision)
    if precision == lib.PREC_FP32_REFINED and FP32_LAYOUT != lib.LAYOUT_AUTO:
        ctx.set_fp32_layout(FP32_LAYOUT)
    if precision == lib.PREC_FP32_REFINED and BLOCK_PRECOND is not None:
        ctx.set_block_precond(BLOCK_PRECOND)
    ctx.set_state(r=eq["r"], v=eq["v"], x=eq["x"])
    ctx.set_thermostat(eq["vq"])
    if predictor is not None:
        ctx.set_predictor(predictor)
    ctx.force()
    ctx.md_begin()
    return ctx


def run_gpu(deck, precision: int, steps: int, warmup: int, eq, predictor=None, kernels=True):
    ctx = start_context(deck, precision, eq, predictor)
    n_save = max(deck.n_save, 1)
    step = 1
    if warmup:
        step += ctx.md_run(step, warmup, n_save)
    t_start = ctx.observables()["temp"]
    # the timed steps run as the product runs them (no per-phase timers, so
    # unobserved steps end without a host synchronisation)
    builds0 = ctx.nb_builds()
    syncs0 = ctx.solve_syncs()
    t0 = time.perf_counter()
    done = ctx.md_run(step, steps, n_save)
    t1 = time.perf_counter()
    assert done == steps, done
    builds = ctx.nb_builds() - builds0
    syncs = ctx.solve_syncs() - syncs0
    step += done
    t_end = ctx.observables()["temp"]
    # the phase breakdown from as many further steps with the library's
    # HIP-event timers on (they synchronise at every step's end)
    ctx.reset_timing()
    ctx.enable_timing(True)
    done = ctx.md_run(step, steps, n_save)
    assert done == steps, done
    t, calls, iters = ctx.timing()
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
        "solve_syncs_per_step": syncs / steps,
        "nb_skin_nm": ctx.skin,
        "warm_start_predictor": ctx.predictor,
        "fp32_layout": {1: "rows", 2: "tiles"}.get(ctx.fp32_layout),
        "block_precond_every": ctx.block_precond if ctx.block_precond >= 0 else None,
        "temp_start_of_timed_window": t_start,
        "temp_end_of_timed_window": t_end,
    }
    if not kernels:
        ctx.close()
        return out
    kern = {}
    klist = [("spmv", lib.KERNEL_SPMV), ("assemble", lib.KERNEL_ASSEMBLE),
             ("pair", lib.KERNEL_PAIR), ("neighbour", lib.KERNEL_NEIGHBOUR),
             ("cpim_mm", lib.KERNEL_CPIM_MM)]
    if precision == lib.PREC_FP32_REFINED:
        klist.append(("apply64", lib.KERNEL_APPLY64))
    klist.append(("solve", lib.KERNEL_SOLVE))       # last: from the last MD step's warm start
    for name, kid in klist:
        # the SpMV, the roofline's kernel, over 100 launches so that the event
        # pair's own overhead is small against them
        kern[name] = ctx.time_kernel(kid, 10 if kid == lib.KERNEL_SOLVE else (100 if kid == lib.KERNEL_SPMV else 20))
    out["kernel_ms"] = kern
    # the pair kernel's work: every entry of every full row (rc + skin) is
    # one distance test, and each pair inside rc is evaluated from both ends
    entries = int(ctx.neighbour_counts().astype(np.int64).sum())
    out["pair_row_entries"] = entries
    out["pair_Gentries_per_s"] = round(entries / (kern["pair"] * 1e-3) / 1e9, 2)
    ctx.close()
    return out


def rel(a, b):
    den = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (den if den > 0 else 1.0)


def error_vs_oracle(deck, eq, steps: int = 20, crits=(1e-8, 1e-14)):
    """Error of one force call of each GPU path against the FP64 oracle
    (OpenMP build) from an identical state and warm start: the state after
    `steps` FP64 MD steps from the equilibrated start, warm start the
    solution of the step before (the reference's warm start).  At the
    reference's stop test rTr < 1e-8 bTb part of each error is the solver
    tolerance itself (two valid solves may stop that far apart, which the
    FP64 path shows too); with every solve taken to rTr < 1e-14 bTb what
    remains is the FP32 relay's own error."""
    from oracle.oracle import Oracle
    c64 = start_context(deck, lib.PREC_FP64, eq)
    n_save = max(deck.n_save, 1)
    try:
        if steps > 1:
            c64.md_run(1, steps - 1, n_save)
        _, _, x_prev = c64.get_state()          # the solution of step `steps` - 1
        c64.md_run(steps, 1, n_save)
        r, v, _ = c64.get_state()               # the positions of step `steps`
    finally:
        c64.close()
    nm, n = deck.n_metal, deck.n_atoms
    out = {"note": "one force call from an identical state and warm start; errors are relative "
                   "2-norms against the FP64 oracle (OpenMP build, every pair: the corrected "
                   "pair count); each entry holds all solves at the stop test rTr < crit bTb; "
                   "at the reference's stop test, *_vs_converged compare with the oracle's "
                   "tight-tolerance solution (the converged answer), for the GPU paths and for "
                   "the oracle's own reference-tolerance solve"}
    converged = None
    for key, crit in sorted(zip(("reference_tolerance", "tight_tolerance"), crits), key=lambda kc: kc[1]):
        o = Oracle(deck, omp=True)
        o.set_solve_crit(crit)
        o.set_state(r, v, x_prev)
        t0 = time.perf_counter()
        rc, fo, po, _, _ = o.force()
        t_or = time.perf_counter() - t0
        _, _, xo, _ = o.get_state()
        it_o = o.solve_stats()[0]
        o.close()
        eo = float(np.sum(po[1:]))
        Fo = fo.reshape(3, n)
        res = {"solve_crit": crit, "oracle_iterations": it_o, "oracle_seconds": round(t_or, 2)}

        def vs_converged(f, pot):
            if converged is None:
                return {}
            fc, ec = converged
            Fc = fc.reshape(3, n)
            F_ = f.reshape(3, n)
            return {"f_vs_converged": float("%.3e" % rel(f, fc)),
                    "f_metal_vs_converged": float("%.3e" % rel(F_[:, :nm], Fc[:, :nm])),
                    "epot_vs_converged": float("%.3e" % (abs(float(np.sum(pot[1:])) - ec) / max(abs(ec), 1e-300)))}
        res["oracle"] = vs_converged(fo, po)
        for name, prec in (("fp32_refined", lib.PREC_FP32_REFINED), ("fp64", lib.PREC_FP64)):
            c = lib.Context(deck, precision=prec)
            try:
                c.set_solve_tolerance(crit)
                c.set_state(r=r, v=v, x=x_prev)
                f, pot, _ = c.force()
                _, _, x = c.get_state()
                it = c.solve_stats()[0]
            finally:
                c.close()
            F = f.reshape(3, n)
            res[name] = {"q": float("%.3e" % rel(x[3 * nm:4 * nm], xo[3 * nm:4 * nm])),
                         "p": float("%.3e" % rel(x[:3 * nm], xo[:3 * nm])),
                         "f": float("%.3e" % rel(f, fo)),
                         "f_metal": float("%.3e" % rel(F[:, :nm], Fo[:, :nm])),
                         "epot": float("%.3e" % (abs(float(np.sum(pot[1:])) - eo) / max(abs(eo), 1e-300))),
                         "iterations": it}
            res[name].update(vs_converged(f, pot))
        if converged is None:
            converged = (fo.copy(), eo)
            del res["oracle"]
        out[key] = res
    return out


def cpu_port(deck, eq, steps: int = 2):
    """The oracle restatement with OpenMP over the reference's own
    decomposition units (i atoms, metal rows, row-groups), every usable core,
    `steps` MD steps from the equilibrated state with its warm start."""
    from oracle.oracle import Oracle
    info = cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", info["usable"] or 1))
    o = Oracle(deck, omp=True)
    o.set_state(eq["r"], eq["v"], eq["x"])
    o.set_thermostat(eq["vq"])
    o.force()                       # the start state's forces (not timed), as capac_md_begin needs
    o.md_begin()
    t_base = o.timers()
    t0 = time.perf_counter()
    for s in range(1, steps + 1):
        o.md_step(s, 10 ** 9)
    sec = (time.perf_counter() - t0) / steps
    t = o.timers() - t_base
    o.close()
    return {"value": 0.0864 / sec, "unit": "ns/day", "cores": threads, "kind": "port",
            "ms_per_step": 1e3 * sec, "solve_ms_per_step": 1e3 * t[7] / steps,
            "nonbonded_ms_per_step": 1e3 * t[4] / steps,
            "nproc": info["nproc"], "cpu_model": info["model"],
            "sample": f"oracle restatement (oracle/cpu_ref.c, OpenMP build, {threads} threads, "
                      f"brute-force pair scan with every pair counted), {steps} MD steps of the same "
                      f"deck from the equilibrated state, warm start the previous solution"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--equil", type=int, default=1500,
                    help="untimed MD steps before the paths start (thermal equilibration)")
    ap.add_argument("--precision", choices=["fp64", "fp32", "both"], default="both")
    ap.add_argument("--fp32-layout", choices=["auto", "rows", "tiles"], default="auto",
                    help="storage layout of the FP32 relay (capac_set_fp32_layout)")
    ap.add_argument("--block-precond", type=int, default=None,
                    help="FP32 path's GMRES preconditioner: block-Jacobi inverses re-formed every N "
                         "assemblies, 0 for the Jacobi diagonal (capac_set_block_precond; default: the "
                         "library's)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the whole box lives on one MI355X)", file=sys.stderr)
        return 2
    global FP32_LAYOUT, BLOCK_PRECOND
    BLOCK_PRECOND = a.block_precond
    FP32_LAYOUT = {"auto": lib.LAYOUT_AUTO, "rows": lib.LAYOUT_ROWS, "tiles": lib.LAYOUT_TILES}[a.fp32_layout]
    deck_spec = gen.config(a.config)
    tmp = tempfile.mkdtemp(prefix="capac_bench_")
    gen.write_deck(deck_spec, tmp)
    deck = lib.Deck.from_dir(tmp)
    log(f"deck {deck_spec.name}: N={deck.n_atoms} Nm={deck.n_metal} n_mat={deck.n_mat}")
    eq = equilibrate(deck, a.equil)

    paths = {}
    precs = {"fp64": lib.PREC_FP64, "fp32": lib.PREC_FP32_REFINED}
    wanted = ["fp64", "fp32"] if a.precision == "both" else [a.precision]
    for name in wanted:
        try:
            paths[name] = run_gpu(deck, precs[name], a.steps, a.warmup, eq)
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
                                                            eq, predictor=0, kernels=False)
            log("fp32 with the previous solution as start", json.dumps(paths["fp32_previous_solution_start"]))
        except lib.CapacError as e:
            log(f"fp32 run with predictor 0 failed: {e}")
    best = "fp32" if "fp32" in paths else "fp64"
    main_path = paths[best]

    nm, n_mat = deck.n_metal, deck.n_mat
    w = 8 if best == "fp64" else 4
    spmv_ms = main_path["kernel_ms"]["spmv"]
    nbytes = spmv_bytes(nm, n_mat, w)
    achieved = nbytes / (spmv_ms * 1e-3) / 1e9
    layout = main_path.get("fp32_layout")
    roofline = {"kernel": "relay SpMV (" + ("FP64 matrix-free relay, one BiCGSTAB half-iteration"
                                             if best == "fp64" else f"FP32 {layout} layout relay, one GMRES column")
                          + "'s product)",
                "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles("spmv_" + best + (f"_{layout}" if best == "fp32" else ""), nbytes),
                "algorithmic_bytes_per_launch": nbytes,
                "algorithmic_bytes_formula": f"{w} (8 N_m^2 - 3 N_m) + 16 n_mat, N_m = {nm}, n_mat = {n_mat}",
                "ms_per_launch": spmv_ms}
    fmt = format_bytes(best, main_path.get("fp32_layout"), nm, n_mat)
    if fmt:
        # what the kernel's format moves per launch, and the fraction of HBM
        # peak on those bytes: the 4-coefficient FP32 relay stores 16 B per
        # ordered pair (ROWS) or per unordered pair (TILES) where the unique
        # FP32 entries take 32 B per unordered pair, so the fraction on the
        # algorithmic bytes can exceed 1
        roofline["format_bytes_per_launch"] = fmt
        roofline["format_frac"] = round(fmt / (spmv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    for name, p in paths.items():
        if "kernel_ms" in p:
            b = spmv_bytes(nm, n_mat, 8 if name == "fp64" else 4)
            p["spmv_GBps"] = round(b / (p["kernel_ms"]["spmv"] * 1e-3) / 1e9, 1)
    pair_valu = pair_valu_roofline(main_path, deck)
    speedup = None
    if "fp32" in paths and "fp64" in paths:
        # the solve alone (the reference's timer category 7), from the last MD
        # step's system and warm start; in MD the pair forces overlap it
        speedup = round(paths["fp64"]["kernel_ms"]["solve"] / paths["fp32"]["kernel_ms"]["solve"], 3)

    err = None
    if not a.no_error and "fp32" in paths:
        try:
            err = error_vs_oracle(deck, eq)
            log("error vs the FP64 oracle:", json.dumps(err))
        except (lib.CapacError, RuntimeError, OSError) as e:
            log(f"error measurement failed: {e}")

    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_port(deck, eq)
            log("cpu baseline:", json.dumps(cpu))
        except (RuntimeError, OSError) as e:
            log(f"cpu baseline failed: {e}")
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
        "data": "synthetic (seeded generator: Mackay icosahedron + SPC/E lattice water), "
                f"equilibrated {eq['steps']} MD steps before timing",
        "config": {"workload": f"{deck_spec.name}: Au{nm} + {deck_spec.n_water} SPC/E waters, "
                               f"{deck_spec.box[0]:g} nm box, {n_mat} CPIM unknowns, MD step dt 1 fs",
                   "atoms": deck.n_atoms, "path": best, "equilibration_steps": eq["steps"],
                   "temp_after_equilibration": round(eq["temp"], 2)},
        "solve_ms_per_step": round(main_path["solve_ms_per_step"], 4),
        "solve_ms_alone": round(main_path["kernel_ms"]["solve"], 4),
        "solve_speedup_fp32_refined_vs_fp64": speedup,
        "temp_timed_window": [round(main_path["temp_start_of_timed_window"], 2),
                              round(main_path["temp_end_of_timed_window"], 2)],
        "error_vs_fp64_oracle": err,
        "paths": paths,
        "roofline": roofline,
        "pair_valu": pair_valu,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
