"""Summaries of rocprofv3 measurements of this library's kernels.

``python -m mountainsort_amd.profiling <kernel_stats.csv> <title> [command] [passes] [note]``
prints a markdown table of the msamd kernels in a rocprofv3
``--stats --output-format csv`` kernel_stats.csv (other kernels, e.g. torch's
synthetic-data generation outside the timed region, are summed into one
line).

``python -m mountainsort_amd.profiling traffic FETCH.db WRITE.db KERNEL COPY_BYTES
PREFIX "fetch command" "write command" [CALLS]`` turns two ``rocprofv3 --pmc
FETCH_SIZE`` / ``--pmc WRITE_SIZE`` results databases into the HBM bytes per
launch of KERNEL, each counter scaled by what it reads for a copy kernel of
known size (an elementwise negation) run in the same program (PREFIX.md and
PREFIX.json).  KERNEL may list the kernels of one stage, comma-separated;
the figures are then per call of the stage, CALLS being its calls in the
program.
"""
import csv
import re
import sys


def _demangle_nested(name):
    """Readable form of an Itanium-mangled nested kernel name that the
    profiler left mangled (it gives up on _Float16 vector parameters, DF16_):
    the qualified name and its integral / bool template arguments, e.g.
    _ZN5msamd12_GLOBAL__N_112apply_kernelILb1ELi6EEEv... -> msamd::apply_kernel<true, 6>.
    Anything else is returned unchanged."""
    if not name.startswith("_ZN"):
        return name
    i, parts = 3, []
    try:
        while i < len(name) and name[i] != "E":
            if name[i].isdigit():
                j = i
                while name[j].isdigit():
                    j += 1
                n = int(name[i:j])
                ident = name[j:j + n]
                i = j + n
                if ident != "_GLOBAL__N_1":
                    parts.append(ident)
            elif name[i] == "I":
                i += 1
                args = []
                while name[i] != "E":
                    if name[i] != "L":
                        return name
                    k = name.index("E", i)
                    lit = name[i + 1:k]
                    if lit[0] == "b":
                        args.append("true" if lit[1:] == "1" else "false")
                    else:
                        v = lit[1:]
                        args.append("-" + v[1:] if v.startswith("n") else v)
                    i = k + 1
                i += 1
                parts[-1] += "<" + ", ".join(args) + ">"
            else:
                return name
    except (IndexError, ValueError):
        return name
    return "::".join(parts) if parts else name


def clean(name):
    name = _demangle_nested(name)
    name = name.replace("(anonymous namespace)::", "")
    name = re.sub(r"^void ", "", name)
    return re.sub(r"\(.*", "", name)


def summarize(path, title, command=None, per=1, note=None):
    rows = list(csv.DictReader(open(path)))
    ours = [r for r in rows if clean(r["Name"]).startswith("msamd::")]
    other_ns = sum(float(r["TotalDurationNs"]) for r in rows
                   if not clean(r["Name"]).startswith("msamd::"))
    total = sum(float(r["TotalDurationNs"]) for r in ours)
    out = [f"# {title}", ""]
    if command:
        out += ["Command:", "", "```", command, "```", ""]
    out += ["| kernel | calls | total ms | avg us | min us | max us | share of msamd time |",
            "|---|---:|---:|---:|---:|---:|---:|"]
    for r in sorted(ours, key=lambda r: -float(r["TotalDurationNs"])):
        t = float(r["TotalDurationNs"])
        out.append(f"| `{clean(r['Name'])}` | {r['Calls']} | {t / 1e6:.3f} | "
                   f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['MinNs']) / 1e3:.2f} | "
                   f"{float(r['MaxNs']) / 1e3:.2f} | {100 * t / total:.1f}% |")
    out += ["", f"msamd kernels total: {total / 1e6:.3f} ms over the profiled run "
            f"({total / 1e6 / per:.3f} ms per pass when the run made {per} passes); "
            f"other kernels (input generation, torch): {other_ns / 1e6:.3f} ms."]
    if note:
        out += ["", note]
    return "\n".join(out) + "\n"




def call_spans(trace_csv, pattern, min_kernels=100):
    """Spans of the library calls made of many kernels whose names contain
    `pattern` (e.g. the filter's three stages over its window batches, on two
    streams): from a `rocprofv3 --kernel-trace` CSV, each maximal run of such
    kernels with no other kernel between them, first start to last end.
    Returns [(span_ms, kernel_count, busy_ms)], busy = summed kernel time."""
    rows = list(csv.DictReader(open(trace_csv)))
    ks = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"])
                for r in rows)
    out, cur = [], None
    for s, e, n in ks + [(0, 0, "")]:
        if pattern and pattern in n:
            if cur is None:
                cur = [s, e, 0, 0]
            cur[1] = max(cur[1], e)
            cur[2] += 1
            cur[3] += e - s
        else:
            if cur is not None and cur[2] >= min_kernels:
                out.append(((cur[1] - cur[0]) / 1e6, cur[2], cur[3] / 1e6))
            cur = None
    return out


def main_spans(argv):
    """profiling spans <trace.csv> <pattern> <out.md> <title> <command>"""
    trace, pattern, out, title, command = argv[:5]
    spans = call_spans(trace, pattern)
    lines = [f"# {title}", "", "Command:", "", "```", command, "```", "",
             f"Calls made of kernels whose names contain `{pattern}`, from the kernel trace "
             "(first start to last end of each run of them):", "",
             "| call | kernels | span ms | summed kernel time ms |", "|---:|---:|---:|---:|"]
    for i, (sp, n, busy) in enumerate(spans):
        lines.append(f"| {i} | {n} | {sp:.3f} | {busy:.3f} |")
    lines.append("")
    open(out, "w").write("\n".join(lines) + "\n")
    sys.stdout.write("\n".join(lines) + "\n")


# --------------------------------------------------------------------------
# PMC traffic: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs (rocpd databases)
# --------------------------------------------------------------------------
def pmc_counters(db_path):
    """{kernel: {counter: [values per dispatch]}, ...} and {kernel: [durations us]}
    from a rocprofv3 results database."""
    import collections
    import sqlite3
    con = sqlite3.connect(db_path)
    cur = con.cursor()
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for name, cn, cv in cur.execute(
            "select name, counter_name, counter_value from pmc_events"):
        agg[clean(name)][cn].append(float(cv))
    dur = collections.defaultdict(list)
    for name, d in cur.execute("select name, duration from kernels"):
        dur[clean(name)].append(d / 1000.0)
    return agg, dur


def traffic_report(fetch_db, write_db, kernel, copy_bytes, command_fetch, command_write,
                   calls=1):
    """HBM bytes per launch of `kernel` from FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch), each scaled by the factor that makes the same counter read the
    known bytes of the calibration kernel (an elementwise negation, whose name
    contains 'neg', reading and writing copy_bytes per launch).  Returns
    (dict, text).  kernel may be a comma-separated list of the kernels of one
    stage; its figures are then per call of the stage (calls of it in the
    measured program)."""
    fa, fdur = pmc_counters(fetch_db)
    wa, _ = pmc_counters(write_db)

    def avg(agg, pred, counter):
        vals = [v for k, c in agg.items() if pred(k) for v in c.get(counter, [])]
        return sum(vals) / len(vals) if vals else None

    def calib(agg, counter):
        # the calibration kernel: an elementwise negation reading and writing
        # copy_bytes, the only negation in the measured program
        vals = [v for k, c in agg.items() if "neg" in k.lower() and "msamd" not in k
                for v in c.get(counter, [])]
        if not vals:
            return None
        calib.used[counter] = len(vals)
        return sum(vals) / len(vals)

    calib.used = {}

    names = kernel.split(",")
    is_k = lambda k: any(k.endswith(n) or k.split("<")[0].endswith(n) for n in names)
    if len(names) > 1:
        # a stage of several kernels: the counters of all their dispatches,
        # per call of the stage (calls = the stage's calls in the program)
        def avg(agg, pred, counter):
            vals = [v for k, c in agg.items() if pred(k) for v in c.get(counter, [])]
            return sum(vals) / calls if vals else None
    f_k, w_k = avg(fa, is_k, "FETCH_SIZE"), avg(wa, is_k, "WRITE_SIZE")
    f_c, w_c = calib(fa, "FETCH_SIZE"), calib(wa, "WRITE_SIZE")
    f_scale = copy_bytes / (f_c * 1024.0) if f_c else None
    w_scale = copy_bytes / (w_c * 1024.0) if w_c else None
    durs = [d for k, v in fdur.items() if is_k(k) for d in v]
    if len(names) > 1 and durs:
        durs = [sum(durs) / calls]
    out = {"kernel": kernel,
           "fetch_kib_per_launch": f_k, "write_kib_per_launch": w_k,
           "copy_bytes": copy_bytes, "copy_fetch_kib": f_c, "copy_write_kib": w_c,
           "fetch_scale": f_scale, "write_scale": w_scale,
           "copy_dispatches_used": dict(calib.used),
           "avg_duration_us_fetch_run": sum(durs) / len(durs) if durs else None}
    if None not in (f_k, w_k, f_scale, w_scale):
        out["read_bytes"] = f_k * 1024.0 * f_scale
        out["write_bytes"] = w_k * 1024.0 * w_scale
        out["traffic_bytes"] = out["read_bytes"] + out["write_bytes"]
    lines = [f"# HBM traffic of {kernel}", "", "Commands:", "", "```", command_fetch,
             command_write, "```", "",
             "| quantity | value |", "|---|---:|"]
    for k, v in out.items():
        lines.append(f"| {k} | {v:.6g} |" if isinstance(v, float) else f"| {k} | {v} |")
    return out, "\n".join(lines) + "\n"


def main_traffic(argv):
    """python -m mountainsort_amd.profiling traffic FETCH.db WRITE.db KERNEL COPY_BYTES
    OUT_PREFIX "fetch command" "write command" -> OUT_PREFIX.md and OUT_PREFIX.json"""
    import json
    fetch_db, write_db, kernel, copy_bytes, prefix, cf, cw = argv[:7]
    calls = int(argv[7]) if len(argv) > 7 else 1
    out, text = traffic_report(fetch_db, write_db, kernel, float(copy_bytes), cf, cw, calls)
    if calls > 1:
        out["stage_calls"] = calls
    open(prefix + ".md", "w").write(text)
    json.dump(out, open(prefix + ".json", "w"), indent=1)
    sys.stdout.write(text)



def main_all(argv):
    """python -m mountainsort_amd.profiling pmc-all FETCH.db WRITE.db OUT.md TITLE
    ["fetch command" "write command"] -> per-kernel average FETCH_SIZE / WRITE_SIZE
    (KiB per dispatch, raw counter values) and durations of every msamd kernel
    in the runs, with the calibration factors of the negation kernel."""
    fetch_db, write_db, out, title = argv[:4]
    cmds = argv[4:6]
    fa, fdur = pmc_counters(fetch_db)
    wa, _ = pmc_counters(write_db)
    lines = [f"# {title}", ""]
    if cmds:
        lines += ["Commands:", "", "```", *cmds, "```", ""]
    calib = {}
    for agg, counter in ((fa, "FETCH_SIZE"), (wa, "WRITE_SIZE")):
        vals = [v for k, c in agg.items() if "neg" in k.lower() and "msamd" not in k
                for v in c.get(counter, [])]
        calib[counter] = sum(vals) / len(vals) if vals else float("nan")
    lines += [f"Calibration: a 1 GiB elementwise negation (reads and writes 1 GiB) reads "
              f"FETCH_SIZE {calib['FETCH_SIZE']:.6g} KiB and WRITE_SIZE "
              f"{calib['WRITE_SIZE']:.6g} KiB, so read bytes = FETCH_SIZE x "
              f"{1048576 / calib['FETCH_SIZE']:.4g} KiB and write bytes = WRITE_SIZE x "
              f"{1048576 / calib['WRITE_SIZE']:.4g} KiB.", ""]
    lines += ["| kernel | dispatches | avg us | FETCH_SIZE KiB | WRITE_SIZE KiB |",
              "|---|---:|---:|---:|---:|"]
    names = sorted((k for k in fdur if "msamd" in k), key=lambda k: -sum(fdur[k]))
    for k in names:
        f = fa.get(k, {}).get("FETCH_SIZE", [])
        w = wa.get(k, {}).get("WRITE_SIZE", [])
        d = fdur[k]
        lines.append(f"| `{k}` | {len(d)} | {sum(d) / len(d):.2f} | "
                     f"{(sum(f) / len(f)) if f else float('nan'):.6g} | "
                     f"{(sum(w) / len(w)) if w else float('nan'):.6g} |")
    open(out, "w").write("\n".join(lines) + "\n")
    sys.stdout.write("\n".join(lines) + "\n")


def main_counters(argv):
    """python -m mountainsort_amd.profiling counters KERNEL OUT.md TITLE "command" DB [DB ...]
    -> for each counter of each results database, its value for KERNEL summed
    over the counter's hardware instances and averaged over the kernel's
    dispatches, with the instance count and the mean dispatch time."""
    import collections
    import sqlite3
    kernel, out, title, command = argv[:4]
    lines = [f"# {title}", "", "Command(s):", "", "```", command, "```", "",
             "| counter | per dispatch (sum over instances) | instances | mean per instance |",
             "|---|---:|---:|---:|"]
    durs = []
    for db in argv[4:]:
        con = sqlite3.connect(db)
        vals = collections.defaultdict(list)
        for name, cn, cv in con.execute("select name, counter_name, counter_value from pmc_events"):
            if clean(name).split("<")[0].endswith(kernel):
                vals[cn].append(float(cv))
        mine = [d / 1000.0 for name, d in con.execute("select name, duration from kernels")
                if clean(name).split("<")[0].endswith(kernel)]
        durs += mine
        nd = max(len(mine), 1)
        for cn, v in sorted(vals.items()):
            inst = len(v) // nd
            lines.append(f"| {cn} | {sum(v) / nd:.4g} | {inst} | {sum(v) / max(len(v), 1):.4g} |")
    if durs:
        lines += ["", f"Mean dispatch time of {kernel}: {sum(durs) / len(durs):.1f} us "
                  f"({len(durs)} dispatches)."]
    open(out, "w").write("\n".join(lines) + "\n")
    sys.stdout.write("\n".join(lines) + "\n")


def csv_counters(path):
    """{kernel: {counter: [value per dispatch]}} and {kernel: [durations us]} from a
    rocprofv3 counter_collection.csv (one row per dispatch and counter)."""
    import collections
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    dur = collections.defaultdict(dict)
    for r in csv.DictReader(open(path)):
        k = clean(r["Kernel_Name"])
        agg[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
        dur[k][r["Dispatch_Id"]] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0
    return agg, {k: list(v.values()) for k, v in dur.items()}


def main_all_csv(argv):
    """python -m mountainsort_amd.profiling pmc-all-csv FETCH.csv WRITE.csv OUT.md TITLE
    ["fetch command" "write command"] -> as pmc-all, from counter_collection.csv files."""
    fetch_csv, write_csv, out, title = argv[:4]
    cmds = argv[4:6]
    fa, fdur = csv_counters(fetch_csv)
    wa, _ = csv_counters(write_csv)
    lines = [f"# {title}", ""]
    if cmds:
        lines += ["Commands:", "", "```", *cmds, "```", ""]
    calib = {}
    for agg, counter in ((fa, "FETCH_SIZE"), (wa, "WRITE_SIZE")):
        vals = [v for k, c in agg.items() if "neg" in k.lower() and "msamd" not in k
                for v in c.get(counter, [])]
        calib[counter] = sum(vals) / len(vals) if vals else float("nan")
    fs, ws = 1048576 / calib["FETCH_SIZE"], 1048576 / calib["WRITE_SIZE"]
    lines += [f"Calibration: a 1 GiB elementwise negation (reads and writes 1 GiB) reads "
              f"FETCH_SIZE {calib['FETCH_SIZE']:.6g} KiB and WRITE_SIZE "
              f"{calib['WRITE_SIZE']:.6g} KiB, so read bytes = FETCH_SIZE x {fs:.4g} KiB "
              f"and write bytes = WRITE_SIZE x {ws:.4g} KiB.", ""]
    lines += ["| kernel | dispatches | avg us | FETCH_SIZE KiB | WRITE_SIZE KiB | read GB | "
              "written GB |", "|---|---:|---:|---:|---:|---:|---:|"]
    names = sorted((k for k in fdur if "msamd" in k), key=lambda k: -sum(fdur[k]))
    for k in names:
        f = fa.get(k, {}).get("FETCH_SIZE", [])
        w = wa.get(k, {}).get("WRITE_SIZE", [])
        d = fdur[k]
        fm = sum(f) / len(f) if f else float("nan")
        wm = sum(w) / len(w) if w else float("nan")
        lines.append(f"| `{k}` | {len(d)} | {sum(d) / len(d):.2f} | {fm:.6g} | {wm:.6g} | "
                     f"{fm * fs * 1024 / 1e9:.3f} | {wm * ws * 1024 / 1e9:.3f} |")
    open(out, "w").write("\n".join(lines) + "\n")
    sys.stdout.write("\n".join(lines) + "\n")


def main_counters_csv(argv):
    """python -m mountainsort_amd.profiling counters-csv REGEX OUT.md TITLE "command"
    CSV [CSV ...] -> for each msamd kernel matching REGEX in rocprofv3
    counter_collection.csv files (one row per dispatch and counter), each
    counter averaged over the kernel's dispatches, the dispatch time, and the
    derived MFMA busy fraction SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE x
    128 SIMDs per XCD) and texture-unit fractions per CU (x / (GRBM_GUI_ACTIVE /
    8 XCDs x 256 CUs)) when those counters are present."""
    import collections
    regex, out, title, command = argv[:4]
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    durs = collections.defaultdict(dict)
    for path in argv[4:]:
        for r in csv.DictReader(open(path)):
            k = clean(r["Kernel_Name"])
            if not k.startswith("msamd::") or not re.search(regex, k):
                continue
            vals[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
            durs[k][(path, r["Dispatch_Id"])] = (int(r["End_Timestamp"]) -
                                                 int(r["Start_Timestamp"])) / 1000.0
    lines = [f"# {title}", "", "Command(s):", "", "```", command, "```", ""]
    for k in sorted(vals):
        d = list(durs[k].values())
        lines += [f"## `{k}`", "", f"Mean dispatch time {sum(d) / len(d):.1f} us over "
                  f"{len(d)} dispatches (counter runs).", "",
                  "| counter | mean per dispatch |", "|---|---:|"]
        mean = {c: sum(v) / len(v) for c, v in vals[k].items()}
        for c in sorted(mean):
            lines.append(f"| {c} | {mean[c]:.4g} |")
        g = mean.get("GRBM_GUI_ACTIVE")
        derived = []
        if g and "SQ_VALU_MFMA_BUSY_CYCLES" in mean:
            # GRBM_GUI_ACTIVE summed over the 8 XCDs; the MFMA busy cycles
            # summed over every SIMD (1024 = 8 x 128)
            derived.append(("MFMA busy (per SIMD)",
                            mean["SQ_VALU_MFMA_BUSY_CYCLES"] / (g * 128)))
        for c in ("TD_TD_BUSY", "TD_TC_STALL", "TA_TA_BUSY", "TA_ADDR_STALLED_BY_TC_CYCLES",
                  "TCP_PENDING_STALL_CYCLES"):
            if c in mean and "GRBM_GUI_ACTIVE" in mean:
                derived.append((f"{c} / cycles (per CU)", mean[c] / (g / 8 * 256)))
        if derived:
            lines += ["", "| derived | fraction |", "|---|---:|"]
            lines += [f"| {n} | {v:.3f} |" for n, v in derived]
        lines.append("")
    open(out, "w").write("\n".join(lines) + "\n")
    sys.stdout.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "counters":
        main_counters(sys.argv[2:])
    elif len(sys.argv) > 1 and sys.argv[1] == "spans":
        main_spans(sys.argv[2:])
    elif len(sys.argv) > 1 and sys.argv[1] == "counters-csv":
        main_counters_csv(sys.argv[2:])
    elif len(sys.argv) > 1 and sys.argv[1] == "pmc-all-csv":
        main_all_csv(sys.argv[2:])
    elif len(sys.argv) > 1 and sys.argv[1] == "traffic":
        main_traffic(sys.argv[2:])
    elif len(sys.argv) > 1 and sys.argv[1] == "pmc-all":
        main_all(sys.argv[2:])
    else:
        cmd = sys.argv[3] if len(sys.argv) > 3 else None
        per = int(sys.argv[4]) if len(sys.argv) > 4 else 1
        note = sys.argv[5] if len(sys.argv) > 5 else None
        sys.stdout.write(summarize(sys.argv[1], sys.argv[2], cmd, per, note))
