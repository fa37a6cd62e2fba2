#!/usr/bin/env python3
"""Benchmark of the 2D TMz FDTD fitness evaluation on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--autotune] [--serve N]

One "step" is one fitness evaluation of the plain path on the 1024^2,
dx = 0.25 mm, 8192-step workload (one antenna; one simulation per launch
sequence), with every input already resident on the device.  `value` is
Mcell-steps/s of the timed steps (cell-step = one Yee cell advanced one full
E+H step), timed without per-launch events; a second pass over the same
evaluations with HIP events on every launch gives the kernel figures.  The
same JSON line carries:
  * batched: one batch of the 512^2 x 4096-step workload, 64 parameter
    vectors x 16 antennas = 1024 simulations advanced together (64 evals),
    and the plain path on the first vectors of the same batch (evals/s ratio);
  * roofline: the dominant update kernel's algorithmic bytes per launch over
    its average launch time, timed with HIP events on the stream it runs on,
    the other update kernels beside it, and the whole update phase; `traffic`
    is the rocprofv3-measured HBM bytes per launch when recorded for the
    same schedule (profiles/update_traffic.json);
  * cpu_baseline: the oracle's FP32 restatement on one host core, on a bounded
    sample of the same 1024^2 workload (a baseline, not the target);
  * serving (with --serve): p50/p99 latency of a continuous-batching run.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import tomography_amd as t  # noqa: E402
from tomography_amd.workloads import (BATCH_SCHEDULE, PLAIN_SCHEDULE, REF_SCHEDULE, SCHED_KEYS, SERVE_ANTENNAS,  # noqa: E402
                                      SERVE_GRIDS, batch_vectors, cell_steps, seeded_vectors, serve_requests,
                                      uneven_requests)

HBM_PEAK_GBS = 8000.0
METRIC = "Mcell-steps/s & fitness evals/s, batched vs plain; p50/p99 eval latency"


def log(msg):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


UPDATE_KERNELS = ("k_tblock", "k_tedge", "k_trow", "k_tcol", "k_tinner", "k_tfused", "k_tres", "k_twave")


def _kernel_roofline(name, k):
    if k["launches"] == 0 or k["ms"] <= 0:
        return None
    avg_s = k["ms"] / k["launches"] / 1e3
    bpl = k["alg_bytes"] / k["launches"]
    achieved = bpl / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": name,
            "launches": k["launches"], "avg_launch_us": round(avg_s * 1e6, 2), "alg_bytes_per_launch": bpl,
            "alg_bytes_per_cell_step": round(k["alg_bytes"] / k["cell_steps"], 3) if k["cell_steps"] else None,
            "cell_steps_per_launch": k["cell_steps"] / k["launches"], "total_ms": round(k["ms"], 2)}


def roofline_from(stats):
    """Roofline of the dominant update kernel (the one that moves the most
    algorithmic bytes; the edge kernels run beside it on their own streams),
    the other update kernels, and the whole update phase (every stream, wall
    time) as one effective rate."""
    ks = [r for r in (_kernel_roofline(n, stats[n]) for n in UPDATE_KERNELS) if r]
    if not ks:
        return None
    dom = dict(max(ks, key=lambda r: r["alg_bytes_per_launch"] * r["launches"]))
    dom["other_kernels"] = [r for r in ks if r["kernel"] != dom["kernel"]]
    u = stats["update"]
    if u["ms"] > 0:
        eff = u["alg_bytes"] / (u["ms"] / 1e3) / 1e9
        dom["update_phase"] = {"achieved": round(eff, 1), "frac": round(eff / HBM_PEAK_GBS, 4),
                               "wall_ms": round(u["ms"], 2), "epochs": u["launches"],
                               "Mcell_steps_per_s": round(u["cell_steps"] / (u["ms"] / 1e3) / 1e6, 1)}
    return dom


def traffic_for(workload, kernel, schedule):
    """Measured HBM bytes per launch of `kernel` (profiles/update_traffic.json:
    rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel
    calibration), if recorded for this workload and schedule."""
    path = os.path.join(ROOT, "profiles", "update_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        rec = json.load(open(path)).get(workload, {})
    except (OSError, ValueError):
        return None
    if tuple(rec.get("schedule", ())) != tuple(schedule):
        return None
    return rec.get("kernels", {}).get(kernel, {}).get("hbm_bytes_per_launch")


def add_traffic(rl, traffic, tsteps):
    """Counter-measured HBM bytes per launch beside the algorithmic figure:
    the real HBM fraction, and the time-blocked minimum (a launch advances
    tsteps steps, so the least it can move is the algorithmic bytes / tsteps,
    the state read and written once)."""
    if not rl:
        return rl
    rl["traffic"] = traffic
    tmin = rl["alg_bytes_per_launch"] / tsteps
    rl["blocked_min_bytes"] = round(tmin, 1)
    if traffic:
        avg_s = rl["avg_launch_us"] / 1e6
        rl["traffic_frac"] = round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4)
        rl["traffic_over_blocked_min"] = round(traffic / tmin, 3)
    return rl


def kernel_ms(st, per=1.0):
    return {n: round(st[n]["ms"] / per, 3) for n in st
            if isinstance(st[n], dict) and n != "update" and st[n]["launches"]}


def apply_schedule(ctx, cfg, n_sims, sched, tune):
    if tune:
        ctx.autotune(cfg, n_sims)
    else:
        ctx.set_schedule(**sched)
    return dict(zip(SCHED_KEYS, ctx.schedule()))


def run_plain(ctx, steps, warmup, sched, tune):
    """cfg2: one 1024^2 simulation per eval, one eval after another."""
    cfg = t.preset(t.PRESET_CFG2)
    used = apply_schedule(ctx, cfg, 1, sched, tune)
    vecs = seeded_vectors(steps + warmup, 8, 20261202)
    for i in range(warmup):
        ctx.eval(cfg, vecs[i])  # the first call also builds and caches the target far field
    t0 = time.perf_counter()
    fits = [ctx.eval(cfg, vecs[warmup + i]) for i in range(steps)]
    dt = time.perf_counter() - t0
    # The same evaluations again with HIP events on every launch (the events
    # ride on the dispatch packets but still lengthen each launch gap, so the
    # timed pass above runs without them): per-kernel launch statistics.
    ctx.profile(True)
    t1 = time.perf_counter()
    fits2 = [ctx.eval(cfg, vecs[warmup + i]) for i in range(steps)]
    dt_prof = time.perf_counter() - t1
    st = ctx.stats()
    ctx.profile(False)
    assert all(np.isfinite(fits)) and fits2 == fits
    return cfg, dt, st, used, dt_prof


def run_batched(ctx, sched, tune):
    """cfg3: 64 vectors x 16 antennas = 1024 simulations advanced together."""
    cfg = t.preset(t.PRESET_CFG3)
    used = apply_schedule(ctx, cfg, 128, sched, tune)
    vecs = batch_vectors()
    ctx.target_farfield(cfg)  # warm-up: the 16-antenna target phantom
    ctx.profile(True)
    t0 = time.perf_counter()
    fits = ctx.eval_batch(cfg, vecs)
    dt = time.perf_counter() - t0
    st = ctx.stats()
    ctx.profile(False)
    assert np.all(np.isfinite(fits))
    return cfg, dt, st, len(vecs), fits, used


def run_plain_same_workload(ctx, n, sched, batch_fits):
    """The plain path on the batched workload: the first n vectors of the cfg3
    batch, each evaluated by tomo_eval (its 16 simulations one at a time)."""
    cfg = t.preset(t.PRESET_CFG3)
    ctx.set_schedule(**sched)
    vecs = batch_vectors()[:n]
    ctx.eval(cfg, vecs[0])  # warm-up (tile lists, caches)
    t0 = time.perf_counter()
    fits = [ctx.eval(cfg, v) for v in vecs]
    dt = time.perf_counter() - t0
    rel = max(abs(f - b) / abs(b) for f, b in zip(fits, batch_fits[:n]))
    return {"evals": n, "s": round(dt, 3), "evals_per_s": round(n / dt, 3),
            "Mcell_steps_per_s": round(cell_steps(cfg) * n / dt / 1e6, 1), "max_rel_fitness_diff_vs_batched": rel}


def run_serve(ctx, n_req, rates, plain_sample, batch_sched, plain_sched, seed=20261205):
    """Continuous batching under Poisson arrivals at each rate: requests are
    submitted at their arrival times while others are resident and retire as
    they finish; p50/p99 latency is submit -> result.  The plain path serves the
    same mix one request at a time (its service time is a lower bound of its
    latency)."""
    reqs = serve_requests(n_req, seed)
    for c in {(r[0].nx, r[0].n_exc) for r in reqs}:  # targets are set-up, not timed
        ctx.target_farfield(t.preset(t.PRESET_CFG5, c[0]).copy(n_exc=c[1]))
    cells = sum(cell_steps(r[0]) for r in reqs)
    out = {"workload": f"{n_req} requests, grids {SERVE_GRIDS[0]}^2..{SERVE_GRIDS[-1]}^2 log-spaced, "
                       f"antennas {SERVE_ANTENNAS}, 8-32-dim vectors, Poisson arrivals (CFG5)",
           "mean_cell_steps_per_request": cells / n_req, "sweep": []}
    ctx.set_schedule(**batch_sched)
    rng = np.random.default_rng(seed + 7)
    for rate in rates:
        arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
        t0 = time.perf_counter()
        pending, lat, done, i, last = {}, [], 0, 0, t0
        while done < n_req:
            now = time.perf_counter() - t0
            while i < n_req and arrivals[i] <= now:
                pending[ctx.submit(reqs[i][0], reqs[i][1])] = True
                i += 1
            if not pending:
                time.sleep(max(0.0, min(0.01, arrivals[i] - now)))
                continue
            ctx.progress()
            for tk in list(pending):
                ok, fit, l = ctx.poll(tk)
                if ok:
                    lat.append(l)
                    del pending[tk]
                    done += 1
            if time.perf_counter() - last > 10:
                log(f"serve rate {rate}: {done}/{n_req} done")
                last = time.perf_counter()
        wall = time.perf_counter() - t0
        out["sweep"].append({"rate_per_s": rate, "wall_s": round(wall, 3), "evals_per_s": round(n_req / wall, 3),
                             "Mcell_steps_per_s": round(cells / wall / 1e6, 1),
                             "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
                             "p99_latency_s": round(float(np.percentile(lat, 99)), 4)})
    if plain_sample:
        ctx.set_schedule(**plain_sched)
        svc, cp = [], 0.0
        for c, v in reqs[:plain_sample]:
            t1 = time.perf_counter()
            ctx.eval(c, v)
            svc.append(time.perf_counter() - t1)
            cp += cell_steps(c)
        out["plain"] = {"requests": plain_sample, "evals_per_s": round(len(svc) / sum(svc), 3),
                        "Mcell_steps_per_s": round(cp / sum(svc) / 1e6, 1),
                        "p50_service_s": round(float(np.percentile(svc, 50)), 4),
                        "p99_service_s": round(float(np.percentile(svc, 99)), 4)}
    return out


def run_search(ctx, method, pop, evals, inflight, sched):
    """The asynchronous search driver (tomo_search_run) on the 512^2,
    4-antenna serving configuration: a closed loop that issues a new request
    as soon as a result returns."""
    cfg = t.preset(t.PRESET_CFG5, 512).copy(n_exc=4)
    ctx.target_farfield(cfg)
    ctx.set_schedule(**sched)
    lo = np.array([-15.0, -15.0, 2.0, 10.0, 10.0, 8.0, 0.0, 0.0])
    hi = np.array([15.0, 15.0, 6.0, 30.0, 50.0, 20.0, 1.0, 0.0])
    best, res = ctx.search(cfg, lo, hi, method, pop=pop, max_evals=evals, inflight=inflight, seed=20261206)
    res = {k: (round(v, 4) if isinstance(v, float) and k != "best_fitness" else v) for k, v in res.items()}
    res["method"] = "DE" if method == t.SEARCH_DE else "PSO"
    res.update({"pop": pop, "inflight": inflight, "workload": "512^2, 4 antennas, 8-dim (CFG5 grid)",
                "Mcell_steps_per_s": round(res["evals"] * cell_steps(cfg) / res["wall_s"] / 1e6, 1),
                "best_x": [round(float(x), 3) for x in best]})
    return res


def run_uneven(ctx, n, plain_sample, batch_sched, plain_sched):
    """All n requests submitted at once and run to completion through the
    request loop (continuous batching over mixed grids), against the plain
    path (tomo_eval, one simulation per launch sequence) on a sample."""
    reqs = uneven_requests(n)
    ctx.set_schedule(**batch_sched)
    for g in sorted({r[0].nx for r in reqs}):  # targets are set-up, not timed
        ctx.target_farfield(t.preset(t.PRESET_CFG4, g))
    ctx.profile(True)
    t0 = time.perf_counter()
    tickets = [ctx.submit(c, v) for c, v in reqs]
    lat, pending = {}, set(tickets)
    last = time.perf_counter()
    while pending:
        ctx.progress()
        if time.perf_counter() - last > 10:
            log(f"uneven: {n - len(pending)}/{n} done")
            last = time.perf_counter()
        for tk in list(pending):
            ok, fit, l = ctx.poll(tk)
            if ok:
                lat[tk] = l
                pending.discard(tk)
    wall = time.perf_counter() - t0
    st = ctx.stats()
    ctx.profile(False)
    cells = sum(cell_steps(c) for c, _ in reqs)
    out = {"workload": f"{n} requests, grids log-uniform 256^2..2048^2, steps clamp(8N, 2048, 16384), 1 antenna, "
                       "0-3 tumours (CFG4)", "cell_steps": cells, "wall_s": round(wall, 3),
           "Mcell_steps_per_s": round(cells / wall / 1e6, 1), "evals_per_s": round(n / wall, 3),
           "p50_latency_s": round(float(np.percentile(list(lat.values()), 50)), 3),
           "p99_latency_s": round(float(np.percentile(list(lat.values()), 99)), 3),
           "update_phase_Mcell_steps_per_s": round(st["update"]["cell_steps"] / (st["update"]["ms"] / 1e3) / 1e6, 1)
           if st["update"]["ms"] else None}
    if plain_sample:
        # The plain path's time on the whole mix, estimated per grid-size bin
        # (bins a factor sqrt(2) wide): `plain_sample` seeded requests of each
        # bin are timed, and the bin's plain seconds per cell-step is applied
        # to all of its requests.  The cost per cell-step depends mainly on
        # the grid, and the bins keep the heavy tail (cost ~ N^3) represented.
        ctx.set_schedule(**plain_sched)
        edges = [256, 362, 512, 724, 1024, 1448, 2049]
        rng = np.random.default_rng(7)
        bins, pred, tt, nt = [], 0.0, 0.0, 0
        for lo, hi in zip(edges[:-1], edges[1:]):
            members = [i for i, (c, _) in enumerate(reqs) if lo <= c.nx < hi]
            if not members:
                continue
            pick = rng.choice(members, min(plain_sample, len(members)), replace=False)
            tp, cp = 0.0, 0.0
            for i in pick:
                c, v = reqs[i]
                t1 = time.perf_counter()
                ctx.eval(c, v)
                tp += time.perf_counter() - t1
                cp += cell_steps(c)
            cb = sum(cell_steps(reqs[i][0]) for i in members)
            pred += cb * tp / cp
            tt += tp
            nt += len(pick)
            bins.append({"nx": [lo, hi], "requests": len(members), "timed": len(pick),
                         "plain_Mcell_steps_per_s": round(cp / tp / 1e6, 1)})
        out["plain_mix_estimate"] = {
            "method": "per grid-size bin: plain seconds per cell-step of the timed requests x the bin's cell-steps",
            "requests_timed": nt, "s_timed": round(tt, 3), "predicted_s": round(pred, 3),
            "evals_per_s": round(n / pred, 3), "Mcell_steps_per_s": round(cells / pred / 1e6, 1), "bins": bins}
        out["batched_vs_plain_evals_per_s"] = round(out["evals_per_s"] / out["plain_mix_estimate"]["evals_per_s"], 3)
    return out


def run_precision(ctx, n, max_grid=320, oracle_grid=288):
    """Per-result error of the Kahan-FP32 running DFT and of the FP64 one
    (both with FP32 fields) against the FP64 oracle, on the smallest grids of
    the uneven set: relative L2 of the far field and relative fitness error."""
    from oracle.binding import Oracle
    o = Oracle()
    reqs = [r for r in uneven_requests(512) if r[0].nx <= max_grid][:n]
    rows = []
    for c, v in reqs:
        log(f"precision: grid {c.nx}, {c.steps} steps")
        row = {"grid": c.nx, "steps": c.steps, "tumour_blocks": len(v) // 8}
        f32, d32, _ = ctx.eval(c, v, want_farfield=True)
        c64 = c.copy(flags=c.flags | t.FLAG_DFT_FP64)
        f64, d64, _ = ctx.eval(c64, v, want_farfield=True)
        row["rel_D_kahan_vs_fp64dft"] = float(np.linalg.norm(d32 - d64) / np.linalg.norm(d64))
        if c.nx <= oracle_grid:
            fo, do, _ = o.eval(c, v, fp64=True)
            row["rel_D_kahan_vs_oracle"] = float(np.linalg.norm(d32 - do) / np.linalg.norm(do))
            row["rel_D_fp64dft_vs_oracle"] = float(np.linalg.norm(d64 - do) / np.linalg.norm(do))
            row["rel_fit_kahan_vs_oracle"] = float(abs(f32 - fo) / abs(fo))
            row["rel_fit_fp64dft_vs_oracle"] = float(abs(f64 - fo) / abs(fo))
        rows.append(row)
    keys = [k for k in rows[0] if k.startswith("rel_")] if rows else []
    summary = {k: float(max(r[k] for r in rows if k in r)) for k in keys if any(k in r for r in rows)}
    return {"results": rows, "max": summary}


def cpu_baseline(steps=2000):
    """The oracle's FP32 restatement on one host core, first `steps` steps of
    the 1024^2 plain workload (fields, CPML, Debye, running DFT)."""
    from oracle.binding import Oracle
    o = Oracle()
    cfg = t.preset(t.PRESET_CFG2)
    v = seeded_vectors(1, 8, 20261202)[0]
    s = o.sim(cfg, v, exc=0, fp64=False)
    t0 = time.perf_counter()
    s.step(steps)
    dt = time.perf_counter() - t0
    mcs = cfg.nx * cfg.ny * steps / dt / 1e6
    return {"value": round(mcs, 3), "unit": "Mcell-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle FP32 (fdtd_oracle.cpp), 1024^2 cfg2 phantom, steps 0-{steps - 1} of 8192, 1 thread; "
                      f"{dt:.1f} s", "evals_per_s": round(mcs * 1e6 / cell_steps(cfg), 6),
            "host": platform_host()}


# One oracle sample in a fresh interpreter that never loads the HIP library:
# argv = repository root, config bytes (hex), parameter vector, steps.
_CPU_WORKER = r"""
import sys, time
sys.path.insert(0, sys.argv[1])
from tomography_amd._abi import SimConfig
from oracle.binding import Oracle
cfg = SimConfig.from_buffer_copy(bytes.fromhex(sys.argv[2]))
v = [float(x) for x in sys.argv[3].split(",")]
s = Oracle().sim(cfg, v, exc=0, fp64=False)
t0 = time.perf_counter()
s.step(int(sys.argv[4]))
print(time.perf_counter() - t0)
"""


def cpu_baseline_all_cores(steps=500, max_procs=16):
    """The same oracle sample in one process per host core (at most
    `max_procs`), all at once: the arrangement of one search worker per core.
    Aggregate Mcell-steps/s over the slowest process's time."""
    nproc = max(1, min(os.cpu_count() or 1, max_procs))
    cfg = t.preset(t.PRESET_CFG2)
    v = seeded_vectors(1, 8, 20261202)[0]
    args = [sys.executable, "-c", _CPU_WORKER, ROOT, bytes(cfg).hex(), ",".join(repr(float(x)) for x in v),
            str(steps)]
    procs = [subprocess.Popen(args, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
             for _ in range(nproc)]
    outs = [p.communicate()[0] for p in procs]
    if any(p.returncode for p in procs):
        return None
    each = [float(o.strip().splitlines()[-1]) for o in outs]
    mcs = nproc * cfg.nx * cfg.ny * steps / max(each) / 1e6
    return {"value": round(mcs, 1), "unit": "Mcell-steps/s", "cores": nproc, "kind": "port",
            "sample": f"oracle FP32, 1024^2 cfg2 phantom, steps 0-{steps - 1}, {nproc} processes at once "
                      f"(one per core), slowest {max(each):.1f} s"}


def run_ref_compat(ctx, sched, n_eval=2, ref_steps=250):
    """The drop-in path: FDTD_GPU() and FDTD_CPU() semantics (401^2, 4
    excitations x 2000 steps, 10 freq x 100 angles) evaluated through the C ABI,
    beside the reference's own FDTD_CPU.cxx compiled from its sources
    (oracle/_ref, when present) on a bounded sample: one excitation, the first
    `ref_steps` steps, extrapolated to a whole evaluation."""
    ctx.set_schedule(**sched)
    out = {"schedule": dict(zip(SCHED_KEYS, ctx.schedule()))}
    rng = np.random.default_rng(20261207)
    for name, pid in (("FDTD_GPU", t.PRESET_REF_GPU), ("FDTD_CPU", t.PRESET_REF_CPU)):
        cfg = t.preset(pid)
        n2 = cfg.numpatches ** 2
        if cfg.flags & t.FLAG_PARAMS_EPS_SIGMA:
            vecs = [np.concatenate([rng.uniform(1, 7, n2), rng.uniform(0, 0.3, n2)]) for _ in range(n_eval + 1)]
        else:
            vecs = [np.concatenate([rng.uniform(1, 7, n2), rng.uniform(1.5, 41, n2), rng.uniform(0, 0.3, n2)])
                    for _ in range(n_eval + 1)]
        # tomo_fitness: the call the FDTD_GPU / FDTD_CPU shims make (the
        # vector's excitations advance together), against the packaged
        # measurement as the shims use it
        ctx.fitness(pid, vecs[0])  # warm-up
        t0 = time.perf_counter()
        fits = [ctx.fitness(pid, v) for v in vecs[1:]]
        dt = (time.perf_counter() - t0) / n_eval
        out[name] = {"s_per_eval": round(dt, 4), "evals_per_s": round(1 / dt, 3),
                     "Mcell_steps_per_s": round(cell_steps(cfg) / dt / 1e6, 1)}
        # the same evaluations with HIP events on every launch: per-kernel
        # launch counts and durations (a cooperative launch is timed by the
        # events around it on its stream)
        ctx.profile(True)
        fits2 = [ctx.fitness(pid, v) for v in vecs[1:]]
        st = ctx.stats()
        ctx.profile(False)
        assert np.all(np.isfinite(fits)) and fits2 == fits, (name, fits, fits2)
        out[name]["kernels_per_eval"] = {
            k: {"launches": st[k]["launches"] / n_eval, "avg_us": round(st[k]["ms"] * 1e3 / st[k]["launches"], 2)}
            for k in st if isinstance(st[k], dict) and k != "update" and st[k]["launches"]}
        out[name]["resident_refused"] = st["resident_refused"]
    try:
        from oracle.binding import Reference
        if Reference.available():
            ref = Reference()
            cfg = t.preset(t.PRESET_REF_CPU)
            v = np.concatenate([rng.uniform(1, 7, 81), rng.uniform(0, 0.3, 81)])
            t0 = time.perf_counter()
            ref.run(v, 0, ref_steps, want_d=False)
            dt = time.perf_counter() - t0
            per_eval = dt * (cfg.n_exc * cfg.steps) / ref_steps
            out["reference_FDTD_CPU"] = {
                "kind": "reference", "cores": 1, "s_per_eval": round(per_eval, 2),
                "Mcell_steps_per_s": round(cfg.nx * cfg.ny * ref_steps / dt / 1e6, 2),
                "sample": f"FDTD_CPU.cxx compiled from the reference sources (oracle/_ref), excitation 0, "
                          f"steps 0-{ref_steps - 1} of 2000, 1 thread; {dt:.1f} s, extrapolated to 4 x 2000 steps"}
            out["FDTD_CPU_gpu_vs_reference_cpu"] = round(per_eval / out["FDTD_CPU"]["s_per_eval"], 1)
    except (ImportError, OSError):
        pass
    return out


REF_MAIN = os.path.join(ROOT, "oracle", "_ref", "fdtd_cpu_main")


def start_reference_full_run():
    """The reference as shipped: the whole FDTD_CPU() call (FDTD_CPU.cxx:669-892,
    compiled unmodified into oracle/_ref/fdtd_cpu_main; 401^2, 4 excitations x
    2000 steps) once on one host core, in its own process, started before the
    GPU legs and timed beside them."""
    if not os.path.exists(REF_MAIN):
        return None
    rng = np.random.default_rng(20261208)
    v = np.concatenate([rng.uniform(1, 7, 81), rng.uniform(0, 0.3, 81)])
    p = subprocess.Popen([REF_MAIN], stdin=subprocess.PIPE, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                         text=True)
    t0 = time.perf_counter()
    p.stdin.write(" ".join(repr(float(x)) for x in v))
    p.stdin.close()
    p.stdin = None  # communicate() later only reads
    return p, t0


def finish_reference_full_run(run, limit_s=420):
    if run is None:
        return None
    p, t0 = run
    try:
        out, _ = p.communicate(timeout=max(1.0, limit_s - (time.perf_counter() - t0)))
    except subprocess.TimeoutExpired:
        p.kill()
        p.communicate()
        return {"status": f"stopped after {limit_s} s"}
    dt = time.perf_counter() - t0
    lines = out.strip().splitlines()
    cfg = t.preset(t.PRESET_REF_CPU)
    return {"kind": "reference", "cores": 1, "s_per_eval": round(dt, 2), "evals_per_s": round(1 / dt, 5),
            "Mcell_steps_per_s": round(cell_steps(cfg) / dt / 1e6, 2),
            "returned": lines[-1].split()[-1] if lines else None,
            "sample": "the whole FDTD_CPU() call as shipped (oracle/_ref/fdtd_cpu_main: 401^2, 4 excitations x "
                      "2000 steps), one process on one host core, run while the GPU legs ran"}


def platform_host():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count()}


def hip_copy_calibration(mb):
    """A plain device-to-device copy of known size (rocprof counter calibration)."""
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    n = mb << 20
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ref", action="store_true", help="skip the reference-compatible (drop-in) leg")
    ap.add_argument("--serve", type=int, default=48, help="requests of the continuous-batching study (per rate; "
                    "10000 in its definition, bounded by default; 0 = skip)")
    ap.add_argument("--rate", type=str, default="3,6", help="Poisson arrival rates (requests/s), comma separated")
    ap.add_argument("--serve-plain-sample", type=int, default=8, help="plain-path requests timed on the serving mix")
    ap.add_argument("--search", type=int, default=0, help="evaluations of the asynchronous DE search driver")
    ap.add_argument("--schedule", type=str, default="",
                    help="plain-path schedule tsteps,seg_rows,edge_waves,inner_waves,chunk,grid_mult,form,edge_rows")
    ap.add_argument("--batch-schedule", type=str, default="", help="batched-path schedule (same fields)")
    ap.add_argument("--autotune", action="store_true", help="pick both schedules with the autotuner instead")
    ap.add_argument("--plain-evals", type=int, default=2,
                    help="plain-path evals on the batched workload (for the batched/plain ratio; 0 = skip)")
    ap.add_argument("--calib-copy-mb", type=int, default=0)
    ap.add_argument("--hbm-budget-gb", type=float, default=0.0,
                    help="device-memory budget of the context (GB; 0: the library's default of 90%% of free memory)")
    ap.add_argument("--uneven", type=int, default=128,
                    help="requests of the uneven-batch study (512 in its definition, bounded by default; 0 = skip)")
    ap.add_argument("--uneven-plain-sample", type=int, default=2,
                    help="plain-path evals timed per grid-size bin of the uneven mix")
    ap.add_argument("--precision", type=int, default=0, help="results of the FP32-Kahan vs FP64 precision study")
    ap.add_argument("--ref-schedule", type=str, default="", help="drop-in leg schedule (same fields as --schedule)")
    ap.add_argument("--only-ref", action="store_true",
                    help="run only the reference-compatible (drop-in) leg and print its record (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: tomography_amd runs on one GPU; --gpus {args.gpus} is not supported "
              "(simulations are independent: more GPUs would each run their own share).", file=sys.stderr)
        return 2
    ref_full = None if (args.no_ref or args.only_ref) else start_reference_full_run()
    ctx = t.Context(hbm_budget=int(args.hbm_budget_gb * 1e9))
    rsched = dict(REF_SCHEDULE)
    if args.ref_schedule:
        rsched = dict(zip(SCHED_KEYS, [int(x) for x in args.ref_schedule.split(",")]))
    if args.only_ref:
        print(json.dumps({"ref_compat": run_ref_compat(ctx, rsched)}))
        ctx.close()
        return 0
    tune = args.autotune
    psched = dict(PLAIN_SCHEDULE)
    if args.schedule:
        psched = dict(zip(SCHED_KEYS, [int(x) for x in args.schedule.split(",")]))
    cfg, dt, st, sched, dt_prof = run_plain(ctx, args.steps, args.warmup, psched, tune)
    cs = cell_steps(cfg) * args.steps
    value = cs / dt / 1e6
    out = {"metric": METRIC, "value": round(value, 1), "unit": "Mcell-steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded phantoms)",
           "config": {"workload": "plain: 1024^2 cells, dx 0.25 mm, 8192 steps, 1 antenna, 8 freq x 180 angles, "
                                  "one simulation per launch sequence",
                      "preset": "CFG2", "schedule": sched,
                      "schedule_source": "autotuned" if tune else "fixed"},
           "evals_per_s": round(args.steps / dt, 3),
           "ms_per_step_profiled_pass": round(dt_prof / args.steps * 1e3, 3)}
    rl = roofline_from(st)
    if rl:
        add_traffic(rl, traffic_for("plain_cfg2", rl["kernel"], tuple(sched.values())), sched["tsteps"])
    out["roofline"] = rl
    out["kernel_ms_per_step"] = kernel_ms(st, args.steps)
    bsch = dict(BATCH_SCHEDULE)
    if args.batch_schedule:
        bsch = dict(zip(SCHED_KEYS, [int(x) for x in args.batch_schedule.split(",")]))
    if not args.no_batched:
        bcfg, bdt, bst, nvec, bfits, bsched = run_batched(ctx, bsch, tune)
        bcs = cell_steps(bcfg) * nvec
        brl = roofline_from(bst)
        if brl:
            add_traffic(brl, traffic_for("batched_cfg3", brl["kernel"], tuple(bsched.values())), bsched["tsteps"])
        out["batched"] = {"workload": "512^2, dx 0.5 mm, 4096 steps; 64 vectors x 16 antennas = 1024 simulations, "
                                      "16 freq x 360 angles (CFG3)",
                          "Mcell_steps_per_s": round(bcs / bdt / 1e6, 1), "evals_per_s": round(nvec / bdt, 3),
                          "s_per_batch": round(bdt, 3), "roofline": brl, "schedule": bsched,
                          "kernel_ms": kernel_ms(bst)}
        if args.plain_evals:
            pl = run_plain_same_workload(ctx, args.plain_evals, sched, bfits)
            out["batched"]["plain_same_workload"] = pl
            out["batched"]["batched_vs_plain_evals_per_s"] = round(out["batched"]["evals_per_s"] / pl["evals_per_s"], 3)
    if args.uneven:
        out["uneven"] = run_uneven(ctx, args.uneven, args.uneven_plain_sample, bsch, sched)
    if args.precision:
        out["precision"] = run_precision(ctx, args.precision)
    if args.serve:
        out["serving"] = run_serve(ctx, args.serve, [float(r) for r in args.rate.split(",")], args.serve_plain_sample,
                                   bsch, sched)
    if args.search:
        out["search"] = run_search(ctx, t.SEARCH_DE, 256, args.search, 64, bsch)
    if args.calib_copy_mb:
        hip_copy_calibration(args.calib_copy_mb)
    if not args.no_ref:
        out["ref_compat"] = run_ref_compat(ctx, rsched)
        full = finish_reference_full_run(ref_full)
        if full:
            out["ref_compat"]["reference_FDTD_CPU_full"] = full
            if "s_per_eval" in full:
                out["ref_compat"]["FDTD_CPU_gpu_vs_reference_cpu_full"] = round(
                    full["s_per_eval"] / out["ref_compat"]["FDTD_CPU"]["s_per_eval"], 1)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        out["cpu_baseline"]["ratio_gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
        allc = cpu_baseline_all_cores()
        if allc:
            allc["ratio_gpu_over_cpu"] = round(value / allc["value"], 1)
            out["cpu_baseline"]["all_cores"] = allc
    ctx.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
