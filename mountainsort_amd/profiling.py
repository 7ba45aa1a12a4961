"""Summaries of rocprofv3 kernel statistics for this library's kernels.

``python -m mountainsort_amd.profiling <kernel_stats.csv> <title> [command]``
prints a markdown table of the msamd kernels in a rocprofv3
``--stats --output-format csv`` kernel_stats.csv (other kernels, e.g. torch's
synthetic-data generation outside the timed region, are summed into one
line).
"""
import csv
import re
import sys


def clean(name):
    name = name.replace("(anonymous namespace)::", "")
    name = re.sub(r"^void ", "", name)
    return re.sub(r"\(.*", "", name)


def summarize(path, title, command=None, per=1):
    rows = list(csv.DictReader(open(path)))
    ours = [r for r in rows if "msamd::" in r["Name"]]
    other_ns = sum(float(r["TotalDurationNs"]) for r in rows if "msamd::" not in r["Name"])
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
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    cmd = sys.argv[3] if len(sys.argv) > 3 else None
    per = int(sys.argv[4]) if len(sys.argv) > 4 else 1
    sys.stdout.write(summarize(sys.argv[1], sys.argv[2], cmd, per))
