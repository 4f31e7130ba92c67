#!/usr/bin/env python3
"""Benchmark of the static k-eigenvalue hot path on one MI355X.

A step is one complete static solve (the reference's cmfd::outer from its
first-call state, src/mod_cmfd.f90:415-509, plus PowDis) of a seeded
synthetic core whose inputs are already resident in HBM.  Metric: k-eff
solves per second of the plain FP64 drop-in, with the FP32+refinement form
timed on the same core beside it; both are checked against the FP64
reference result (k within 1 pcm, node powers within 1e-4).

Also reported:
  roofline      BiCGSTAB KA (previous x/r update + p update + 7-point SpMV +
                (rs,v) dot), 17 algorithmic words per node per launch, timed
                with start/stop events attached to one KA dispatch per group
                solve in a profiled solve; traffic from the committed
                rocprofv3 counter summary of the same kernel;
  cpu_baseline  the reference's own Fortran build (oracle/_ref) on one host
                core, on a bounded sample of the same core (outer and nodal
                update times measured free of the deck parse), with the
                host's CPU model and core count.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""

from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "core3d_2g": dict(seed=1, fixture="core3d_2g_s1_solve",
                      desc="3D 2-group quarter core, 34x34x26 nodes (17x17 assemblies of 21.5 cm, 2x2 nodes, "
                           "24 fuel planes of 15 cm + 2 reflector planes), SANM, per-node +-2% XS"),
    "core3d_8g": dict(seed=2, fixture="core3d_8g_s2_solve",
                      desc="3D 8-group quarter core, 68x68x96 nodes (17x17 assemblies, 4x4 nodes, 24 layers x 4), "
                           "SANM, downscatter + upscatter into groups 6-8"),
}


BATCH_SEED = 3


def batch_bench(args):
    """The batching workload: 256 uneven cores (2D/3D, 2-4 groups, 5k-60k
    nodes) in one context per group count, the contexts driven concurrently
    (one host thread and one stream each).  A step solves all 256 cores from
    the first-call state; value = converged solves per second (cores the
    reference itself stops as nodally unstable stop the same way and are not
    counted), FP32+refinement (the workload's form), with FP64 beside it."""
    import threading
    from adpres_amd.configs import batch_core
    from adpres_amd.engine import ADG_OK, Batch
    t0 = time.perf_counter()
    cores = [batch_core(BATCH_SEED, i).problem for i in range(256)]
    log(f"bench: batch256 generated in {time.perf_counter() - t0:.1f} s: {sum(q.nnod for q in cores)} nodes, "
        f"{sum(q.unknowns for q in cores)} unknowns")
    fx = os.path.join(ROOT, "tests", "golden", f"batch_s{BATCH_SEED}_solve.npz")
    ref = np.load(fx) if os.path.exists(fx) else None
    classes = {}
    for i, q in enumerate(cores):
        classes.setdefault(q.ng, []).append(i)

    def run(precision):
        bats = {G: Batch([cores[i] for i in idx], precision) for G, idx in classes.items()}
        tunes = {G: b.autotune() for G, b in bats.items()}
        log(f"bench: batch256 {precision} autotune {tunes}")
        keff = np.zeros(len(cores))
        niter = np.zeros(len(cores), np.int32)
        status = np.zeros(len(cores), np.int32)

        def solve(G):
            bats[G].reset()
            st, ni, ke = bats[G].outer()
            keff[classes[G]] = ke
            niter[classes[G]] = ni
            status[classes[G]] = st

        def step():
            th = [threading.Thread(target=solve, args=(G,)) for G in bats]
            for t in th:
                t.start()
            for t in th:
                t.join()

        for _ in range(args.warmup):
            step()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            step()
        dt = (time.perf_counter() - t1) / args.steps
        keff_t, status_t, niter_t = keff.copy(), status.copy(), niter.copy()
        # roofline of KA over a whole context: HIP events on one steady-state
        # KA launch per group solve, in the first outers of a fresh solve
        # (every core still iterating, so each launch covers every core)
        wsz = 4 if precision == "fp32_refine" else 8
        roof_ctx, tot_b, tot_t = {}, 0.0, 0.0
        for G, b in bats.items():
            nodes = int(sum(cores[i].nnod for i in classes[G]))
            b.set_profiling(True)
            b.reset()
            b.outer(nout=6)
            pr = b.read_profile()
            b.set_profiling(False)
            if pr["k1_n"]:
                nb = 17 * wsz * nodes
                roof_ctx[str(G)] = {"nodes": nodes, "bytes_per_launch": nb, "avg_launch_us": round(pr["k1_ms"] * 1e3, 3),
                                    "GBps": round(nb / (pr["k1_ms"] * 1e-3) / 1e9, 1), "samples": pr["k1_n"]}
                tot_b += nb * pr["k1_n"]
                tot_t += pr["k1_ms"] * 1e-3 * pr["k1_n"]
        keff[:], status[:], niter[:] = keff_t, status_t, niter_t
        for b in bats.values():
            b.close()
        solved = int(np.sum(status == ADG_OK))
        roof = None
        if tot_t > 0:
            ach = tot_b / tot_t / 1e9
            # measured HBM traffic per launch of the context the counter
            # profile covers (the one with the most nodes)
            kname = f"void adg::k_bicgA<{'float' if wsz == 4 else 'double'}, false, 1>"
            traffic, _ = measured_traffic("batch256", kname)
            Gp = max(classes, key=lambda G: sum(cores[i].nnod for i in classes[G]))
            if traffic and str(Gp) in roof_ctx:
                c = roof_ctx[str(Gp)]
                c["traffic"] = traffic
                c["traffic_GBps"] = round(traffic / (c["avg_launch_us"] * 1e-6) / 1e9, 1)
            roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "traffic_source": f"profiles/pmc_bicg_batch256.txt: per launch over the {Gp}-group context "
                                      "(rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1 KiB, calibrated on a 256 MiB copy)",
                    "kernel": f"k_bicgA<{'float' if wsz == 4 else 'double'}> launched over every core of a context "
                              "(17 words per node)",
                    "contexts": roof_ctx}
        out = {"seconds_per_step": dt, "solves_per_s": solved / dt, "cores_solved": solved, "roofline": roof,
               "cells_per_block": {str(G): t["shape"] for G, t in tunes.items()},
               "cores_stopped": int(len(cores) - solved), "max_outers": int(niter.max())}
        if ref is not None:
            m = np.isfinite(ref["Ke"])
            both = m & (status == ADG_OK)
            out["k_err_pcm_max"] = float(np.max(np.abs(keff[both] - ref["Ke"][both])) * 1e5)
            out["reference_cores_solved"] = int(m.sum())
            out["outcome_differs_from_reference"] = [int(i) for i in np.nonzero(m != (status == ADG_OK))[0]]
        log(f"bench: batch256 {precision}: {out}")
        return out

    r32 = run("fp32_refine")
    r64 = run("fp64") if not args.no_fp32 else None
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_batch([batch_core(BATCH_SEED, i) for i in range(256)])
        log(f"bench: batch256 cpu baseline {cpu}")
    out = {
        "metric": "k-eff solves/s (256-core batch, FP32+refine; FP64 beside it) at 1 pcm",
        "value": round(r32["solves_per_s"], 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(r32["seconds_per_step"] * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (FP64 defect correction)",
        "data": "synthetic (seeded generator, adpres_amd/configs.py batch_core)",
        "config": {"workload": "batch256", "cores": len(cores), "nodes": int(sum(q.nnod for q in cores)),
                   "unknowns": int(sum(q.unknowns for q in cores)),
                   "contexts": {str(G): len(idx) for G, idx in classes.items()}},
        "roofline": r32["roofline"], "cpu_baseline": cpu, "fp32_refine": r32, "fp64": r64}
    print(json.dumps(out))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def lu_flops(m):
    """Exact flop count of the reference's no-pivot LU_solve of order m
    (src/mod_nodal.f90:829-897): elimination, forward and back substitution."""
    return (m - 1) * m * (2 * m - 1) // 3 + 3 * m * (m - 1) // 2 + (m * m - 1) + (m * m + m - 1)


def nodal_flops(p):
    """LU flops of one nodal update: per direction u, N - L_u interface
    systems of order 2G and N + 2 L_u node and boundary systems of order G
    (L_u lines along u); assembly and transcendentals are not counted."""
    G, N = p.ng, p.nnod
    lines = [len(set(zip(p.iy, p.iz))), len(set(zip(p.ix, p.iz))), len(set(zip(p.ix, p.iy)))]
    return sum((N - L) * lu_flops(2 * G) + (N + 2 * L) * lu_flops(G) for L in lines)


def time_solves(eng, steps, warmup):
    """Warm-up, then `steps` timed solves on the production path (graph
    replay, no profiling events); then one more solve with the engine's
    event timers on, for the per-kernel roofline figure."""
    for _ in range(warmup):
        eng.reset()
        eng.outer()
    eng.set_profiling(False)
    ks, niters = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.reset()
        eng.outer()
        pw = eng.powdis()          # synchronises the stream
        ks.append(eng.Ke)
        niters.append(eng.niter)
    dt = time.perf_counter() - t0
    eng.set_profiling(True)
    eng.reset()
    eng.outer()
    pr = eng.read_profile()
    eng.set_profiling(False)
    return dict(seconds=dt, solves_per_s=steps / dt, k=ks[-1], niter=niters[-1], pow=pw,
                k1_avg_ms=pr["k1_ms"] if pr["k1_n"] else None, k1_samples=pr["k1_n"],
                gsolve_avg_ms=pr["gsolve_ms"], nodal_avg_ms=pr["nodal_ms"])


def xs_update_probe(eng, p, workload, seed, reps=10):
    """The device cross-section update (adg_update_xs: XS_updt with boron,
    temperature, density and rod feedback) on the workload's geometry with
    synthetic feedback tables: its kernel time from HIP events and the
    host-inclusive time of the call (parameter upload, kernel, status read)."""
    from adpres_amd.configs import branch_tables, feedback_tables, generate
    gc = generate(workload, seed)
    G = p.ng

    def timed(par):
        for _ in range(3):
            eng.update_xs(*par)
        kern, wall = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            eng.update_xs(*par)
            wall.append(time.perf_counter() - t0)
            kern.append(eng.last_timing_ms()[0])
        return float(np.mean(kern)) * 1e3, float(np.mean(wall)) * 1e3

    def roof(kernel, per_node, us, wall_ms):
        nbytes = per_node * p.nnod
        ach = nbytes / (us * 1e-6) / 1e9
        return {"bound": "hbm", "kernel": kernel, "algorithmic_bytes_per_launch": nbytes,
                "traffic": measured_traffic(workload, f"void adg::{kernel}<{G}>")[0],
                "bytes_per_node": per_node, "avg_launch_us": round(us, 2), "achieved": round(ach, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "call_ms_with_upload": round(wall_ms, 3), "samples": reps,
                "timing": "HIP events around the kernel on the engine stream (adg_last_timing after adg_update_xs)"}
    t, par = feedback_tables(gc)
    eng.set_xs_tables(t)
    # per node: material, rod column, cell index and three parameters in;
    # D, sigr, nuf, sigf (G each) and sigs (G^2) out
    out = roof("k_xs_update_lds", 4 + 4 + 8 + 3 * 8 + 8 * (4 * G + G * G), *timed(par))
    # branch tables (three points on each of the four parameters, rodded
    # sets): the same reads, and the discontinuity factors (6G) out too
    bt, bpar = branch_tables(gc)
    eng.set_xtab_tables(bt)
    us, wall = timed(bpar)
    E = 4 * G + G * G + 6 * G
    br = roof("k_xtab_tile", 4 + 4 + 8 + 3 * 8 + 8 * E, us, wall)
    # the interpolation reads 16 table values per entry (eight corners at two
    # moderator points; rodded nodes twice) from the cache-resident tables:
    # that, not HBM, bounds this kernel
    tb = 8 * 16 * E * p.nnod
    br["table_read_bytes_per_launch"] = tb
    br["table_read_GBs"] = round(tb / (us * 1e-6) / 1e9, 1)
    br["note"] = ("bound by the table reads through L1/L2 (16 per entry, rodded nodes not counted), "
                  "not by HBM; frac is against HBM for the algorithmic bytes only")
    out["branch_tables"] = br
    return out


def host_cpu():
    """CPU model and the number of host cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model, ncpu


def ref_solve(gc, it, workdir, kern=None):
    """One run of the reference build (oracle/_ref/ref_dump: ADPRES's Fortran,
    amdflang -O3) on the deck of gc with %ITER it = (nout, nin, serc, ferc,
    nac, nupd); returns (wall seconds incl. deck parse, the reference's own
    (CMFD, nodal) CPU-time buckets or None when it stopped before printing
    them, whether it converged)."""
    from adpres_amd.problem import write_deck, write_xs_override
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_dump")
    os.makedirs(workdir, exist_ok=True)
    write_deck(os.path.join(workdir, "deck"), kern=kern or gc.problem.kern, iter_card=it, **gc.deck)
    args = [ref, "deck", "solve", "out.bin"]
    if gc.perturbed:
        write_xs_override(os.path.join(workdir, "xs.bin"), gc.problem)
        args.append("xs.bin")
    t0 = time.perf_counter()
    r = subprocess.run(["bash", "-c", "ulimit -s unlimited; exec " + " ".join(args)], cwd=workdir,
                       capture_output=True, text=True)
    wall = time.perf_counter() - t0
    buckets = None
    for line in r.stdout.splitlines():
        if line.strip().startswith("REF_TIMES"):
            f = line.split()
            buckets = (float(f[1]), float(f[2]))
    return wall, buckets, buckets is not None


def cpu_baseline(problem, outers_per_solve, nupd):
    """The reference build on one host core on the same core.  Small cores:
    the whole solve with the same %ITER, timed by the reference's own CPU-time
    buckets (CMFD + nodal update, no deck parse).  Large cores: three bounded
    runs at once on three host cores, each stopped at NOUT by the reference's
    max-outer stop -- A: 2 outers, B: 4 outers (no nodal update in either),
    C: 2 outers with a nodal update at the second -- give the time of one
    outer, (B - A) / 2, and of one nodal update, C - A, free of the deck
    parse; the solve is then the GPU solve's outers and nodal updates at
    those costs.  (The reference refuses NUPD >= NOUT, so A and B run the
    FDM kernel -- the same CMFD outer without nodal updates -- and C, the
    deck's kernel, 3 outers with NUPD = 2: its nodal update is
    C - A - t_outer.)"""
    from adpres_amd.configs import generate
    if not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "ref_dump")):
        return None
    import threading
    model, ncpu = host_cpu()
    gc = generate(problem.name, WORKLOADS[problem.name]["seed"])
    p = gc.problem
    d = tempfile.mkdtemp(prefix="adg_cpu_")
    try:
        if p.nnod * p.ng <= 200_000:
            wall, bk, conv = ref_solve(gc, (p.nout, p.nin, p.serc, p.ferc, p.nac, p.nupd), os.path.join(d, "full"))
            if not conv:
                return None
            solve_s = bk[0] + bk[1]
            return {"value": 1.0 / solve_s, "unit": "solves/s", "cores": 1, "kind": "reference",
                    "sample": "one whole solve of the same core by ADPRES's Fortran (amdflang -O3, same %ITER), "
                              "timed by its own CPU-time buckets",
                    "cpu_model": model, "host_cores": ncpu, "cmfd_s": bk[0], "nodal_s": bk[1],
                    "wall_incl_parse_s": round(wall, 3), "method": "reference buckets fdm_time + nod_time"}
        runs = {"A": ((2, p.nin, 1e-30, 1e-30, p.nac, 1), "FDM"), "B": ((4, p.nin, 1e-30, 1e-30, p.nac, 1), "FDM"),
                "C": ((3, p.nin, 1e-30, 1e-30, p.nac, 2), None)}
        walls = {}

        def go(tag):
            walls[tag] = ref_solve(gc, runs[tag][0], os.path.join(d, tag), kern=runs[tag][1])[0]
        th = [threading.Thread(target=go, args=(t,)) for t in runs]
        for t in th:
            t.start()
        for t in th:
            t.join()
        t_outer = (walls["B"] - walls["A"]) / 2.0
        t_nodal = walls["C"] - walls["A"] - t_outer
        if not t_outer > 0.0:
            return None
        nnod_upd = outers_per_solve // nupd
        solve_s = outers_per_solve * t_outer + nnod_upd * max(t_nodal, 0.0)
        return {"value": 1.0 / solve_s, "unit": "solves/s", "cores": 1, "kind": "reference",
                "sample": f"ADPRES Fortran (amdflang -O3) on the same core: runs of 2 and 4 FDM outers and of 3 "
                          f"outers with a nodal update (walls {walls['A']:.1f}, {walls['B']:.1f}, {walls['C']:.1f} s, "
                          f"concurrently on three host cores), composed into the GPU solve's "
                          f"{outers_per_solve} outers and {nnod_upd} nodal updates",
                "cpu_model": model, "host_cores": ncpu, "t_outer_s": round(t_outer, 4),
                "t_nodal_s": round(t_nodal, 4), "outers": outers_per_solve, "nodal_updates": nnod_upd,
                "method": "t_outer = (wall(4 FDM outers) - wall(2 FDM outers)) / 2, t_nodal = wall(3 outers with "
                          "a nodal update at the 2nd) - wall(2 FDM outers) - t_outer; solve = outers * t_outer + "
                          "nodal_updates * t_nodal"}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def cpu_baseline_batch(cores_gc, sample_every=8, workers=None):
    """The batching workload's CPU baseline: the reference build solving a
    bounded sample of the 256 cores (every sample_every-th) whole, one
    process per host core (at most 16), aggregate converged solves per
    wall-second of that pool."""
    if not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "ref_dump")):
        return None
    from concurrent.futures import ThreadPoolExecutor
    model, ncpu = host_cpu()
    nw = workers or max(1, min(16, ncpu))
    idx = list(range(0, len(cores_gc), sample_every))
    d = tempfile.mkdtemp(prefix="adg_cpu_batch_")
    try:
        def one(i):
            q = cores_gc[i].problem
            return ref_solve(cores_gc[i], (q.nout, q.nin, q.serc, q.ferc, q.nac, q.nupd), os.path.join(d, str(i)))
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=nw) as ex:
            res = list(ex.map(one, idx))
        wall = time.perf_counter() - t0
        conv = sum(1 for r in res if r[2])
        return {"value": conv / wall, "unit": "solves/s", "cores": nw, "kind": "reference",
                "sample": f"{len(idx)} of the {len(cores_gc)} cores (every {sample_every}th) solved whole by "
                          f"ADPRES's Fortran (amdflang -O3), {nw} processes at once on {nw} host cores; "
                          f"{conv} converged (the rest stop as the reference does) in {wall:.1f} s wall",
                "cpu_model": model, "host_cores": ncpu,
                "cmfd_plus_nodal_s_per_solve": round(float(np.mean([r[1][0] + r[1][1] for r in res if r[2]])), 3)
                if conv else None}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def dropin_leg(workload, seed):
    """The Fortran drop-in (adpres_amd/fortran/_build/adpres_gpu: ADPRES's
    unchanged host linked with the GPU hot path) on the workload's deck with
    its %ITER: one run from the deck, as a user runs it.  The solve phase is
    the host's own CMFD + nodal-update buckets (filled from the engine's
    device timers of the outer call); the wall time adds the deck parse, the
    host's cross-section processing, context creation, the upload and the
    printed outputs."""
    import re
    from adpres_amd.configs import generate
    from adpres_amd.problem import write_deck
    exe = os.path.join(ROOT, "adpres_amd", "fortran", "_build", "adpres_gpu")
    if not os.path.exists(exe):
        return None
    gc = generate(workload, seed)
    if gc.perturbed:
        return None    # per-node perturbations are not expressible in a deck
    p = gc.problem
    d = tempfile.mkdtemp(prefix="adg_dropin_")
    try:
        write_deck(os.path.join(d, "deck"), kern=p.kern, iter_card=(p.nout, p.nin, p.serc, p.ferc, p.nac, p.nupd),
                   **gc.deck)
        t0 = time.perf_counter()
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {exe} deck"], cwd=d, capture_output=True,
                           text=True)
        wall = time.perf_counter() - t0
        if r.returncode != 0 or "ADPRES EXIT NORMALLY" not in r.stdout:
            log(f"bench: drop-in run failed: {r.stdout[-800:]} {r.stderr[-800:]}")
            return None
        text = open(os.path.join(d, "deck.out")).read()

        def bucket(name):
            m = re.search(name + r"\s*:\s*([0-9.]+)", text)
            return float(m.group(1)) if m else None
        cmfd, nod = bucket("CMFD time"), bucket("Nodal update time")
        k = float(re.search(r"MULTIPLICATION EFFECTIVE \(K-EFF\) =\s+([0-9.]+)", text).group(1))
        outers = len(re.findall(r"^\s*\d+\s+\d+\.\d{6}\s+\S+E[+-]\d\d\s+\S+E[+-]\d\d\s*$", text, re.M))
        return {"solves_per_s": round(1.0 / (cmfd + nod), 4) if cmfd and nod is not None else None,
                "cmfd_s": cmfd, "nodal_s": nod, "wall_s": round(wall, 2), "k": k, "outers": outers,
                "input_s": bucket("Input reading time"), "xsec_s": bucket("XSEC processing time"),
                "how": "ADPRES's host with the GPU hot path on the workload's deck (no profiling: the outer is "
                       "replayed as a captured graph); solves/s from its CMFD + nodal-update buckets"}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def measured_traffic(workload, kernel):
    """HBM-side bytes per launch of `kernel` from the committed counter
    summaries profiles/pmc_bicg_<workload>.txt and (the flat BiCGSTAB form)
    profiles/pmc_bicg_flat_<workload>.txt (FETCH_SIZE and WRITE_SIZE runs of
    --kernel-probe, scaled by the 256 MiB copy calibration), with the source
    to cite: (bytes, source) or (None, None)."""
    for name in (f"pmc_bicg_{workload}.txt", f"pmc_bicg_flat_{workload}.txt"):
        path = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(path):
            continue
        for line in open(path):
            if line.startswith(kernel + " "):
                return (int(round(float(line[len(kernel):].split()[3]) * 1e6)),
                        f"profiles/{name} (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1 KiB per dispatch, calibrated on a "
                        f"256 MiB copy)")
    return None, None


def copy_calibration():
    """Two device copies of 256 MiB: the kernel of known size the traffic
    counters are calibrated on."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nb = 256 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(nb)) != 0:
            raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nb))
    for _ in range(2):
        hip.hipMemcpy(b, a, ctypes.c_size_t(nb), 3)
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    return nb


def kernel_probe_batch():
    """Hardware-counter program for the batching workload: the copy
    calibration, then one outer iteration of the context holding the most
    nodes of the 256-core batch (every core iterating, so each BiCGSTAB
    launch covers every core of the context), in FP64 and in FP32."""
    from adpres_amd.configs import batch_core
    from adpres_amd.engine import Batch
    nb = copy_calibration()
    cores = [batch_core(BATCH_SEED, i).problem for i in range(256)]
    classes = {}
    for q in cores:
        classes.setdefault(q.ng, []).append(q)
    G, qs = max(classes.items(), key=lambda kv: sum(q.nnod for q in kv[1]))
    nodes = int(sum(q.nnod for q in qs))
    out = {"probe": "kernel-counters-batch", "copy_bytes": nb, "groups": G, "cores": len(qs), "nodes": nodes,
           "ka_algorithmic_bytes": {"fp64": 17 * 8 * nodes, "fp32_refine": 17 * 4 * nodes}}
    for prec in ("fp64", "fp32_refine"):
        b = Batch(qs, prec)
        b.outer(nout=1)
        b.close()
    print(json.dumps(out))


def kernel_probe(p):
    """Short program for hardware-counter runs (rocprofv3 --pmc): two device
    copies of 256 MiB (the calibration kernel of known size), then one
    BiCGSTAB call (group 1, nin=20, from a flat guess) in each precision; on
    the 8-group core also one device cross-section update with feedback
    tables and one with branch tables."""
    import ctypes
    from adpres_amd.engine import Engine
    hip = ctypes.CDLL("libamdhip64.so")
    nb = 256 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(nb)) != 0:
            raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nb))
    for _ in range(2):
        hip.hipMemcpy(b, a, ctypes.c_size_t(nb), 3)
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    rng = np.random.default_rng(0)
    rhs = rng.uniform(0.5, 1.5, p.nnod)
    x0 = np.ones(p.nnod)
    out = {"probe": "kernel-counters", "copy_bytes": nb, "nodes": p.nnod}
    for prec in ("fp64", "fp32_refine"):
        e = Engine(p, prec)
        x = e.bicg(0, 20, rhs, x0)
        out[prec + "_xmax"] = float(np.abs(x).max())
        if prec == "fp64" and p.name == "core3d_8g":
            # one cross-section update with feedback tables and one with
            # branch tables on the same core
            from adpres_amd.configs import branch_tables, feedback_tables, generate
            gc = generate("core3d_8g", WORKLOADS["core3d_8g"]["seed"])
            t, par = feedback_tables(gc)
            e.set_xs_tables(t)
            e.update_xs(*par)
            bt, bpar = branch_tables(gc)
            e.set_xtab_tables(bt)
            e.update_xs(*bpar)
        e.close()
    print(json.dumps(out))


def nodal_probe(p):
    """Short program for hardware-counter runs of the nodal update: 20 FDM
    outers from the first-call state, then two SANM nodal updates (each a
    leakage pass, the node pass, the interface pass, the boundary-face pass
    and the argmax) on that state; in the FP64 form, then in the FP32+refine
    form (the same FP64 update unless ADG_NOD_LU32=1)."""
    from adpres_amd.engine import Engine
    out = {"probe": "nodal-counters", "nodes": p.nnod, "groups": p.ng}
    for prec in ("fp64", "fp32_refine"):
        e = Engine(p, prec)
        e.outer(nout=20, kern="FDM", raise_on_error=False)
        for k in range(2):
            nd, loc = e.nodal_update()
            out[f"{prec}_ndmax_{k}"] = float(nd)
        e.close()
    print(json.dumps(out))


def grad_leg():
    """Differentiation form on its workload, the 34x34x26-node 2-group core
    (210k parameters): the full static solve converged to 1e-10, the
    discrete adjoint, and the gradients of k and of the peak-power objective
    (log-sum-exp, temperature 50) of the full nodal model, with the frozen
    D-hat gradient and one transpose of the nodal update beside them."""
    from adpres_amd.configs import generate
    from adpres_amd.diff import CoreSensitivity
    beta = 50.0
    sens = CoreSensitivity(generate("core3d_2g", 1).problem, model="full")
    e = sens.engine
    x0 = sens.params()
    t0 = time.perf_counter()
    k0, p0 = sens.solve(x0)
    t_solve = time.perf_counter() - t0
    t0 = time.perf_counter()
    e.adjoint(**sens.adj_iter)
    sens._adjoint_ok = True
    t_adj = time.perf_counter() - t0
    t0 = time.perf_counter()
    sens.gradient(1.0, None)
    t_gk, steps_k = time.perf_counter() - t0, e.nodal_adjoint_steps
    nf = int((p0 > 0).sum())
    r = beta * p0 * nf
    wgt = np.exp(r - r.max())
    dJ = wgt / wgt.sum() * nf
    t0 = time.perf_counter()
    sens.gradient(0.0, dJ)
    t_gJ, steps_J = time.perf_counter() - t0, e.nodal_adjoint_steps
    t0 = time.perf_counter()
    e.backward(1.0, None, **sens.adj_iter)
    t_frz = time.perf_counter() - t0
    w = np.ones((6, sens.G, sens.N))
    e.nodal_vjp(w)
    t0 = time.perf_counter()
    for _ in range(10):
        e.nodal_vjp(w)
    t_vjp = (time.perf_counter() - t0) / 10
    out = {"workload": "core3d_2g", "nodes": sens.N, "parameters": int(x0.size), "k": round(k0, 10),
           "solve_s": round(t_solve, 3), "adjoint_s": round(t_adj, 3),
           "grad_k_full_s": round(t_gk, 3), "fixed_point_steps_k": steps_k,
           "grad_objective_full_s": round(t_gJ, 3), "fixed_point_steps_objective": steps_J,
           "grad_k_frozen_ms": round(t_frz * 1e3, 3), "nodal_transpose_ms": round(t_vjp * 1e3, 3),
           "note": "full: adg_backward_nodal (per fixed-point step one transpose of the nodal update and 16 "
                   "generalized-adjoint outers continued across steps, tol 1e-9); frozen: adg_backward (D-hat "
                   "held fixed, with the adjoint already solved); the transpose time includes the host copies of "
                   "w and of the cotangents"}
    sens.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="core3d_8g", choices=sorted(WORKLOADS) + ["batch256"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-dropin", action="store_true", help="skip the Fortran drop-in leg")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernels directly instead of replaying the captured outer (for profilers)")
    ap.add_argument("--kernel-probe", action="store_true",
                    help="run only a copy of known size and one BiCGSTAB call per precision (for counter runs)")
    ap.add_argument("--nodal-probe", action="store_true",
                    help="run only two nodal updates of the workload core (for counter runs)")
    ap.add_argument("--no-grad", action="store_true", help="skip the differentiation leg")
    ap.add_argument("--grad-probe", action="store_true",
                    help="run only the differentiation leg (for profilers)")
    args = ap.parse_args()
    if args.no_graph:
        os.environ["ADG_GRAPH"] = "0"
    if args.gpus != 1:
        print(f"bench.py: this engine runs one core per GPU and this build measures one GPU only; "
              f"--gpus {args.gpus} is not supported (multi-GPU sharding is a design note)", file=sys.stderr)
        sys.exit(2)

    if args.grad_probe:
        print(json.dumps(grad_leg()))
        return
    if args.workload == "batch256":
        if args.kernel_probe:
            kernel_probe_batch()
        else:
            batch_bench(args)
        return

    from adpres_amd.configs import generate
    from adpres_amd.engine import Engine

    w = WORKLOADS[args.workload]
    p = generate(args.workload, w["seed"]).problem
    fx = os.path.join(ROOT, "tests", "golden", w["fixture"] + ".npz")
    ref = np.load(fx) if os.path.exists(fx) else None
    if ref is not None and "p_ix" not in ref.files:
        ref = dict(Ke=ref["Ke"], pow=ref["pow"].astype(np.float64))
    log(f"bench: {args.workload} N={p.nnod} G={p.ng} unknowns={p.unknowns}")
    if args.kernel_probe:
        kernel_probe(p)
        return
    if args.nodal_probe:
        nodal_probe(p)
        return

    eng = Engine(p, "fp64")
    tune64 = eng.autotune()
    info64 = eng.solver_info()
    log(f"bench: autotune fp64 {tune64} {info64}")
    r64 = time_solves(eng, args.steps, args.warmup)
    # the boundary takes host arrays: one cross-section upload (what a caller
    # that changes the cross sections between solves adds per solve)
    t0 = time.perf_counter()
    eng.set_xs(p)
    up_s = time.perf_counter() - t0
    xs_dev = xs_update_probe(eng, p, args.workload, w["seed"])
    eng.close()
    log(f"bench: fp64 {r64['solves_per_s']:.3f} solves/s k={r64['k']:.8f} outers={r64['niter']}")
    r32 = None
    if not args.no_fp32:
        eng = Engine(p, "fp32_refine")
        tune32 = eng.autotune()
        info32 = eng.solver_info()
        log(f"bench: autotune fp32 {tune32} {info32}")
        r32 = time_solves(eng, args.steps, args.warmup)
        eng.close()
        log(f"bench: fp32_refine {r32['solves_per_s']:.3f} solves/s k={r32['k']:.8f} outers={r32['niter']}")

    def errs(r):
        if ref is None or r is None:
            return None
        m = ref["pow"] > 0
        return dict(k_err_pcm=abs(r["k"] - float(ref["Ke"])) * 1e5,
                    pow_max_rel_err=float(np.max(np.abs(r["pow"][m] - ref["pow"][m]) / ref["pow"][m])))

    roof = None
    if r64["k1_avg_ms"] and info64["persistent"]:
        # the persistent BiCGSTAB kernel: one launch per group solve; canonical
        # count 11 words of setup + 31 words per iteration per node
        kc = info64["nodes_per_thread"]
        k1_bytes = (11 + 31 * p.nin) * 8 * p.nnod
        ach = k1_bytes / (r64["k1_avg_ms"] * 1e-3) / 1e9
        traffic, tsrc = measured_traffic(args.workload, f"void adg::k_bicg_persist<double, {kc}>")
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": tsrc,
                "kernel": f"k_bicg_persist<double, {kc}> (one launch per BiCGSTAB group solve: setup and all "
                          f"{p.nin} iterations, matrix and vectors kept in registers, brick halos exchanged)",
                "algorithmic_bytes_per_launch": k1_bytes,
                "algorithmic_count": f"(11 + 31 x nin) words x 8 B x {p.nnod} nodes (canonical BiCGSTAB count: "
                                     f"11 words of setup and 31 per iteration per node)",
                "avg_launch_us": round(r64["k1_avg_ms"] * 1e3, 3), "samples": r64["k1_samples"],
                "timing": "HIP events attached to every persistent launch (hipExtLaunchKernelGGL start/stop events) "
                          "on the engine stream, in a profiled solve after the timed ones",
                "per_iteration_us": round(r64["k1_avg_ms"] * 1e3 / p.nin, 3),
                "solver": info64,
                "note": "the matrix and vectors stay on chip across the iterations, so the canonical bytes over the "
                        "kernel time can exceed the HBM peak; traffic is what crosses HBM per launch",
                "group_solve_avg_us": round(r64["gsolve_avg_ms"] * 1e3, 2),
                "nodal_update_avg_us": round(r64["nodal_avg_ms"] * 1e3, 2)}
    elif r64["k1_avg_ms"]:
        # roofline of the BiCGSTAB stencil kernel KA (k_bicgA): r, p, v, t, x, rs and
        # 7 matrix coefficients in; r, p, v, x out = 17 words per node per launch
        k1_bytes = 17 * 8 * p.nnod
        ach = k1_bytes / (r64["k1_avg_ms"] * 1e-3) / 1e9
        traffic, tsrc = measured_traffic(args.workload, "void adg::k_bicgA<double, false, 0>")
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": tsrc,
                "kernel": "k_bicgA<double> (BiCGSTAB: previous x/r update + p update + 7-point SpMV + (rs,v))",
                "algorithmic_bytes_per_launch": k1_bytes, "avg_launch_us": round(r64["k1_avg_ms"] * 1e3, 3),
                "samples": r64["k1_samples"], "timing": "HIP events attached to one steady-state KA dispatch per group "
                                                        "solve (hipExtLaunchKernelGGL start/stop events), on the engine "
                                                        "stream, in a profiled solve after the timed ones",
                "group_solve_avg_us": round(r64["gsolve_avg_ms"] * 1e3, 2),
                "nodal_update_avg_us": round(r64["nodal_avg_ms"] * 1e3, 2)}
    if roof is not None:
        roof["xs_update"] = xs_dev
        if r64["nodal_avg_ms"]:
            nf = nodal_flops(p)
            gf = nf / (r64["nodal_avg_ms"] * 1e-3) / 1e9
            roof["nodal"] = {"bound": "valu", "flops_per_update": nf, "avg_update_us": round(r64["nodal_avg_ms"] * 1e3, 2),
                             "achieved": round(gf, 1), "peak": 78600.0, "unit": "GFLOP/s (FP64 vector)",
                             "frac": round(gf / 78600.0, 5),
                             "note": "LU_solve flops only (no-pivot LU of order 2G per interface, G per node and "
                                     "boundary); assembly, transcendentals and the D-hat epilogue not counted"}
    dropin = None
    if not args.no_dropin:
        dropin = dropin_leg(args.workload, w["seed"])
        log(f"bench: drop-in {dropin}")
    grad = None
    if not args.no_grad:
        grad = grad_leg()
        log(f"bench: differentiation {grad}")
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(p, r64["niter"], p.nupd)
        log(f"bench: cpu baseline {cpu}")

    out = {
        "metric": "k-eff solves/s (FP64 GPU drop-in; FP32+refine beside it) at 1 pcm",
        "value": round(r64["solves_per_s"], 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r64["seconds"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, adpres_amd/configs.py)",
        "config": {"workload": args.workload, "description": w["desc"], "nodes": p.nnod, "groups": p.ng,
                   "unknowns": p.unknowns, "kern": p.kern,
                   "iter": {"nout": p.nout, "nin": p.nin, "serc": p.serc, "ferc": p.ferc, "nac": p.nac,
                            "nupd": p.nupd},
                   "outers_per_solve": r64["niter"], "launch_shape_fp64": tune64},
        "parity_fp64": errs(r64),
        "with_xs_upload": {"set_xs_ms": round(up_s * 1e3, 2),
                           "solves_per_s": round(1.0 / (r64["seconds"] / args.steps + up_s), 4)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "dropin": dropin,
        "differentiation": grad,
    }
    if r32 is not None:
        out["fp32_refine"] = {"solves_per_s": round(r32["solves_per_s"], 4),
                              "speedup_vs_fp64": round(r32["solves_per_s"] / r64["solves_per_s"], 3),
                              "outers_per_solve": r32["niter"], "parity": errs(r32), "launch_shape": tune32,
                              "solver": info32,
                              "k1_avg_us": round(r32["k1_avg_ms"] * 1e3, 3) if r32["k1_avg_ms"] else None}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
