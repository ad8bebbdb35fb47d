"""Summaries of rocprofv3 runs of the benchmark, as committed under profiles/.

    python -m gtopen_amd.profiling kernels TRACE_DIR ITERATIONS
    python -m gtopen_amd.profiling traffic FETCH_DIR WRITE_DIR COPY_FETCH_DIR COPY_WRITE_DIR ITERATIONS COPY_MIB

`kernels` reads the kernel-stats CSV of a `rocprofv3 --kernel-trace --stats
--output-format csv` run and prints a markdown table (calls, total ms, ms
per iteration, mean microseconds per launch). `traffic` reads the counter
CSVs of `--pmc FETCH_SIZE` and `--pmc WRITE_SIZE` runs of the benchmark and
of a device-to-device copy of COPY_MIB mebibytes (`bench.py --copy-probe`),
scales each counter so the copy reads and writes exactly its size, and
prints JSON: HBM bytes read and written per iteration for every kernel.
"""
import csv
import glob
import json
import os
import re
import sys


def _csvs(root, suffix):
    out = sorted(glob.glob(os.path.join(root, "**", f"*{suffix}"), recursive=True))
    if not out:
        raise FileNotFoundError(f"no *{suffix} under {root}")
    return out


def short_name(name):
    """Kernel name without the namespace, argument list and template
    keywords: 'gto::k_up_action<0, 0>(...)' -> 'k_up_action<0, 0>'."""
    n = name.strip().strip('"')
    depth, end = 0, len(n)
    for k, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            end = k
            break
    n = n[:end].replace("gto::", "")
    return re.sub(r"\s+", " ", n).strip()


def kernel_stats(trace_dir):
    """{kernel: (calls, total ns)} summed over the stats CSVs of a run."""
    out = {}
    for path in _csvs(trace_dir, "kernel_stats.csv"):
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                k = short_name(row["Name"])
                c, t = out.get(k, (0, 0.0))
                out[k] = (c + int(row["Calls"]), t + float(row["TotalDurationNs"]))
    return out


def counter_sums(pmc_dir, counter):
    """{kernel: summed counter value} over every dispatch of a --pmc run."""
    out = {}
    for path in _csvs(pmc_dir, "counter_collection.csv"):
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                k = short_name(row["Kernel_Name"])
                out[k] = out.get(k, 0.0) + float(row["Counter_Value"])
    return out


def kernels_markdown(trace_dir, iterations):
    st = kernel_stats(trace_dir)
    lines = ["| kernel | calls | total ms | ms / iteration | avg µs / launch |", "|---|---:|---:|---:|---:|"]
    for k, (c, t) in sorted(st.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"| `{k}` | {c} | {t / 1e6:.2f} | {t / 1e6 / iterations:.3f} | {t / 1e3 / c:.1f} |")
    return "\n".join(lines)


def traffic(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, iterations, copy_mib):
    """Bytes read / written per iteration per kernel. FETCH_SIZE and
    WRITE_SIZE count KiB; each is scaled by what it reads for the copy of a
    known size (so the copy's own traffic calibrates the counters)."""
    copy_bytes = float(copy_mib) * (1 << 20)

    def copy_kib(d, counter):
        vals = counter_sums(d, counter)
        return max(vals.values())

    fscale = copy_bytes / (copy_kib(copy_fetch_dir, "FETCH_SIZE") * 1024.0)
    wscale = copy_bytes / (copy_kib(copy_write_dir, "WRITE_SIZE") * 1024.0)
    fetch = counter_sums(fetch_dir, "FETCH_SIZE")
    write = counter_sums(write_dir, "WRITE_SIZE")
    ks = {}
    for k in sorted(set(fetch) | set(write)):
        ks[k] = {"read_bytes_per_iter": fetch.get(k, 0.0) * 1024.0 * fscale / iterations,
                 "write_bytes_per_iter": write.get(k, 0.0) * 1024.0 * wscale / iterations}
    return {"iterations_in_run": iterations, "fetch_scale": fscale, "write_scale": wscale, "kernels": ks}


def main(argv):
    if argv and argv[0] == "kernels" and len(argv) == 3:
        print(kernels_markdown(argv[1], int(argv[2])))
        return 0
    if argv and argv[0] == "traffic" and len(argv) == 7:
        print(json.dumps(traffic(argv[1], argv[2], argv[3], argv[4], int(argv[5]), float(argv[6])), indent=1))
        return 0
    print(__doc__, file=sys.stderr)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
