"""Reduce rocprofv3 CSV output to the per-kernel figures the library reports.

    python -m tomography_amd.rocprof_summary stats DIR PREFIX
        per-kernel calls / total / average duration from PREFIX_kernel_stats.csv
    python -m tomography_amd.rocprof_summary pmc DIR PREFIX COUNTER
        per-kernel, per-template average of COUNTER (KiB per dispatch) from
        PREFIX_counter_collection.csv, plus the copy-kernel dispatches used
        for calibration
    python -m tomography_amd.rocprof_summary traffic FETCH.json WRITE.json BENCH.json COPY_MB COMMAND...
        profiles/update_traffic.json from the two `pmc` summaries (FETCH_SIZE,
        WRITE_SIZE) and the profiled run's bench line, scaled by the copy of
        COPY_MB MiB the run made (bench.py --calib-copy-mb)
    python -m tomography_amd.rocprof_summary db FILE.db [KERNEL=LAST ...]
        the same per-kernel figures as `trace` from rocprofv3's default
        SQLite output (PREFIX_results.db)

Kernel names are shortened to `k_name<T>` (the fused step count) so the two
bench workloads, which run different T, stay apart.
"""
import csv
import json
import os
import re
import sys
from collections import defaultdict

_NAME = re.compile(r"(k_[a-z0-9_]+)(?:<([^>]*)>)?")


def short_name(full):
    """'void tomo::(anonymous namespace)::k_tinner<4>(...)' -> 'k_tinner<4>'."""
    head = full.split("(")[0] if "::" not in full.split("(")[0] else full
    for m in _NAME.finditer(head):
        return m.group(1) + (f"<{m.group(2)}>" if m.group(2) else "")
    return full.split("(")[0].strip()[:60]


def stats(d, prefix):
    rows = list(csv.DictReader(open(os.path.join(d, f"{prefix}_kernel_stats.csv"))))
    out = {}
    for r in rows:
        n = short_name(r["Name"])
        out[n] = {"calls": int(r["Calls"]), "total_ms": float(r["TotalDurationNs"]) / 1e6,
                  "avg_us": float(r["AverageNs"]) / 1e3, "min_us": float(r["MinNs"]) / 1e3,
                  "max_us": float(r["MaxNs"]) / 1e3, "pct": float(r["Percentage"])}
    return out


def _last(spec):
    out = {}
    for a in spec:
        k, n = a.split("=")
        out[k] = int(n)
    return out


def trace(d, prefix, spec=()):
    last = _last(spec)
    per = defaultdict(list)
    info = {}
    for r in csv.DictReader(open(os.path.join(d, f"{prefix}_kernel_trace.csv"))):
        n = short_name(r["Kernel_Name"])
        per[n].append((int(r["Dispatch_Id"]), int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
        info[n] = {"vgpr": int(r["VGPR_Count"]), "agpr": int(r["Accum_VGPR_Count"]), "sgpr": int(r["SGPR_Count"]),
                   "lds": int(r["LDS_Block_Size"]), "scratch": int(r["Scratch_Size"]),
                   "wg": int(r["Workgroup_Size_X"]), "grid": int(r["Grid_Size_X"])}
    out = {}
    for n, v in sorted(per.items(), key=lambda kv: -sum(x[1] for x in kv[1])):
        v.sort()
        rec = {"calls": len(v), "total_ms": sum(x[1] for x in v) / 1e6, "avg_us": sum(x[1] for x in v) / len(v) / 1e3}
        if n in last:
            w = v[-last[n]:]
            rec["last"] = len(w)
            rec["avg_us_last"] = sum(x[1] for x in w) / len(w) / 1e3
        rec.update(info[n])
        out[n] = rec
    return out


def pmc(d, prefix, counter, spec=()):
    last = _last(spec)
    per = defaultdict(list)
    copies = []
    for r in csv.DictReader(open(os.path.join(d, f"{prefix}_counter_collection.csv"))):
        if r["Counter_Name"] != counter:
            continue
        v = float(r["Counter_Value"])
        n = short_name(r["Kernel_Name"])
        if "copy" in r["Kernel_Name"].lower():
            copies.append({"kernel": r["Kernel_Name"][:80], "grid": int(r["Grid_Size"]), "kib": v})
        per[n].append((int(r["Dispatch_Id"]), v))
    kern = {}
    for n, v in sorted(per.items()):
        v.sort()
        rec = {"dispatches": len(v), "avg_kib": sum(x[1] for x in v) / len(v)}
        if n in last:
            w = v[-last[n]:]
            rec["last"] = len(w)
            rec["avg_kib_last"] = sum(x[1] for x in w) / len(w)
        kern[n] = rec
    return {"counter": counter, "unit": "KiB per dispatch", "kernels": kern,
            "largest_copy_dispatches": sorted(copies, key=lambda c: -c["kib"])[:3]}


def db(path, spec=()):
    """Per-kernel dispatch statistics from a rocprofv3 results database."""
    import sqlite3
    last = _last(spec)
    con = sqlite3.connect(path)
    per = defaultdict(list)
    info = {}
    q = ("select name, dispatch_id, duration, vgpr_count, accum_vgpr_count, sgpr_count, lds_size, scratch_size, "
         "workgroup_x, grid_x from kernels")
    for name, did, dur, vg, ag, sg, lds, scr, wg, grid in con.execute(q):
        n = short_name(name)
        per[n].append((int(did), int(dur)))
        info[n] = {"vgpr": vg, "agpr": ag, "sgpr": sg, "lds": lds, "scratch": scr, "wg": wg, "grid": grid}
    out = {}
    for n, v in sorted(per.items(), key=lambda kv: -sum(x[1] for x in kv[1])):
        v.sort()
        rec = {"calls": len(v), "total_ms": sum(x[1] for x in v) / 1e6, "avg_us": sum(x[1] for x in v) / len(v) / 1e3}
        if n in last:
            w = v[-last[n]:]
            rec["last"] = len(w)
            rec["avg_us_last"] = sum(x[1] for x in w) / len(w) / 1e3
        rec.update(info[n])
        out[n] = rec
    return out


def traffic(fetch_path, write_path, bench_path, copy_mb, command):
    """HBM bytes per launch of the update kernels (profiles/update_traffic.json)
    from a FETCH_SIZE and a WRITE_SIZE summary (`pmc` output, each with the
    timed region's dispatches) and the bench line of the profiled command.
    Each counter is scaled by what it reads for the largest copy dispatch of
    the run (bench.py --calib-copy-mb copy_mb: a copy of copy_mb MiB)."""
    fetch, write = json.load(open(fetch_path)), json.load(open(write_path))
    lines = [l for l in open(bench_path) if l.lstrip().startswith("{")]
    bench = json.loads(lines[-1])
    copy_bytes = int(copy_mb) * 1024 * 1024
    fcopy, wcopy = fetch["largest_copy_dispatches"][0]["kib"], write["largest_copy_dispatches"][0]["kib"]
    ff, wf = copy_bytes / 1024.0 / fcopy, copy_bytes / 1024.0 / wcopy
    out = {"calibration": {"copy_bytes": copy_bytes, "fetch_kib": fcopy, "write_kib": wcopy,
                           "fetch_factor": ff, "write_factor": wf},
           "command": command}
    sched = lambda s: [s[k] for k in ("tsteps", "seg_rows", "edge_waves", "inner_waves", "chunk", "grid_mult",
                                      "form", "edge_rows", "inner_rows")]
    legs = [("plain_cfg2", bench["config"]["schedule"], bench["roofline"])]
    if bench.get("batched"):
        legs.append(("batched_cfg3", bench["batched"]["schedule"], bench["batched"]["roofline"]))
    for name, s, rl in legs:
        T = s["tsteps"]
        kern = {}
        for r in [rl] + list(rl.get("other_kernels", [])):
            key = f"{r['kernel']}<{T}>"
            fk = fetch["kernels"].get(key, {}).get("avg_kib_last")
            wk = write["kernels"].get(key, {}).get("avg_kib_last")
            if fk is None or wk is None:
                continue
            kern[r["kernel"]] = {"fetch_kib": fk, "write_kib": wk,
                                 "hbm_bytes_per_launch": (fk * ff + wk * wf) * 1024.0,
                                 "alg_bytes_per_launch": r["alg_bytes_per_launch"]}
        out[name] = {"schedule": sched(s), "kernels": kern}
    return out


def main(argv):
    if len(argv) >= 6 and argv[0] == "traffic":
        print(json.dumps(traffic(argv[1], argv[2], argv[3], argv[4], " ".join(argv[5:])), indent=1))
        return 0
    if len(argv) >= 3 and argv[0] == "stats":
        print(json.dumps(stats(argv[1], argv[2]), indent=1))
    elif len(argv) >= 3 and argv[0] == "trace":
        print(json.dumps(trace(argv[1], argv[2], argv[3:]), indent=1))
    elif len(argv) >= 2 and argv[0] == "db":
        print(json.dumps(db(argv[1], argv[2:]), indent=1))
    elif len(argv) >= 4 and argv[0] == "pmc":
        print(json.dumps(pmc(argv[1], argv[2], argv[3], argv[4:]), indent=1))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
