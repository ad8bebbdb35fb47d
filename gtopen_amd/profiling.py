"""Summaries of rocprofv3 runs of the benchmark, as committed under profiles/.

    python -m gtopen_amd.profiling kernels TRACE_DIR ITERATIONS
    python -m gtopen_amd.profiling traffic FETCH_DIR WRITE_DIR COPY_FETCH_DIR COPY_WRITE_DIR ITERATIONS COPY_MIB [WORKLOAD]
    python -m gtopen_amd.profiling report PREFIX ITERATIONS COPY_MIB ALGO_BYTES TITLE BENCH_COMMAND
    python -m gtopen_amd.profiling counters PMC_DIR ITERATIONS

`counters` prints every counter of a `--pmc` run per kernel and iteration
(the kernels with the most SQ_WAVE_CYCLES first), with the LDS
bank-conflict share SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE when both are
present. `report` reads PREFIX_trace, PREFIX_fetch, PREFIX_write and the copy runs
copy_fetch / copy_write beside them, and prints the markdown summary: the
commands, the kernel table, bytes per kernel and the iteration's total
against ALGO_BYTES (the algorithmic bytes of one iteration).

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
    """Kernel name without the return type, namespace and argument list:
    'void gto::k_up_action<0, 0, false>(...)' -> 'k_up_action<0, 0, false>'."""
    n = name.strip().strip('"')
    if n.startswith("void "):
        n = n[5:]
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


def report(prefix, iterations, copy_mib, algo_bytes, title, bench):
    root = os.path.dirname(prefix)
    cf, cw = os.path.join(root, "copy_fetch"), os.path.join(root, "copy_write")
    t = traffic(prefix + "_fetch", prefix + "_write", cf, cw, iterations, copy_mib)
    fk = max(counter_sums(cf, "FETCH_SIZE").values())
    wk = max(counter_sums(cw, "WRITE_SIZE").values())
    out = [f"# {title}", "", "Raw data: `rocprofv3` on one MI355X (gfx950). Commands (repo root, `TMPDIR=/tmp`;",
           "OUT is the output folder):", "", "```"]
    name = os.path.basename(prefix)
    out.append(f"rocprofv3 --kernel-trace --stats --output-format csv -d OUT/{name}_trace -o run -- {bench}")
    out.append(f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/{name}_fetch -o run -- {bench}")
    out.append(f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT/{name}_write -o run -- {bench}")
    out.append(f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/copy_fetch -o copy -- python3 bench.py --copy-probe {copy_mib:g}")
    out.append(f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT/copy_write -o copy -- python3 bench.py --copy-probe {copy_mib:g}")
    out.append(f"python3 -m gtopen_amd.profiling report OUT/{name} {iterations} {copy_mib:g} {algo_bytes:g} '<title>' '<bench command>'")
    out += ["```", "", f"The process runs {iterations} DCFR iterations (warm-up, timed hipGraph replays, and "
            "event-timed eager launches for the per-kernel split); every figure below is per iteration.", "",
            "## Kernel trace (`--kernel-trace --stats`)", "", kernels_markdown(prefix + "_trace", iterations), "",
            "## HBM bytes (`--pmc FETCH_SIZE`, then `--pmc WRITE_SIZE`)", "",
            f"Calibration: a {copy_mib:g} MiB device-to-device copy reads FETCH_SIZE = {fk:.1f} KiB and "
            f"WRITE_SIZE = {wk:.1f} KiB, so the counters are scaled by {t['fetch_scale']:.3f} (fetch) and "
            f"{t['write_scale']:.3f} (write).", "",
            "| kernel | read GB / iteration | written GB / iteration |", "|---|---:|---:|"]
    tr = tw = 0.0
    for k, v in sorted(t["kernels"].items(), key=lambda kv: -(kv[1]["read_bytes_per_iter"] + kv[1]["write_bytes_per_iter"])):
        r, w = v["read_bytes_per_iter"], v["write_bytes_per_iter"]
        tr, tw = tr + r, tw + w
        out.append(f"| `{k}` | {r / 1e9:.3f} | {w / 1e9:.3f} |")
    out += [f"| **iteration** | **{tr / 1e9:.3f}** | **{tw / 1e9:.3f}** |", "",
            f"Whole iteration: {(tr + tw) / 1e9:.2f} GB against {algo_bytes / 1e9:.2f} GB of algorithmic arena "
            f"bytes ({(tr + tw) / algo_bytes:.2f}x)."]
    return "\n".join(out)


def counters_markdown(pmc_dir, iterations):
    names = set()
    for path in _csvs(pmc_dir, "counter_collection.csv"):
        with open(path, newline="") as f:
            names.update(row["Counter_Name"] for row in csv.DictReader(f))
    names = sorted(names)
    vals = {c: counter_sums(pmc_dir, c) for c in names}
    kernels = set().union(*(v.keys() for v in vals.values()))
    order = sorted(kernels, key=lambda k: -vals.get("SQ_WAVE_CYCLES", {}).get(k, 0.0))
    share = "SQ_LDS_BANK_CONFLICT" in vals and "SQ_LDS_IDX_ACTIVE" in vals
    head = ["kernel"] + names + (["bank-conflict share"] if share else [])
    lines = ["| " + " | ".join(head) + " |", "|---" + "|---:" * (len(head) - 1) + "|"]
    for k in order:
        row = [f"`{k}`"] + [f"{vals[c].get(k, 0.0) / iterations:.4g}" for c in names]
        if share:
            act = vals["SQ_LDS_IDX_ACTIVE"].get(k, 0.0)
            row.append(f"{vals['SQ_LDS_BANK_CONFLICT'].get(k, 0.0) / act:.3f}" if act else "-")
        lines.append("| " + " | ".join(row) + " |")
    return "\n".join(lines)


def main(argv):
    if argv and argv[0] == "kernels" and len(argv) == 3:
        print(kernels_markdown(argv[1], int(argv[2])))
        return 0
    if argv and argv[0] == "traffic" and len(argv) in (7, 8):
        t = traffic(argv[1], argv[2], argv[3], argv[4], int(argv[5]), float(argv[6]))
        if len(argv) == 8:
            t = {"workload": argv[7], **t}
        print(json.dumps(t, indent=1))
        return 0
    if argv and argv[0] == "counters" and len(argv) == 3:
        print(counters_markdown(argv[1], int(argv[2])))
        return 0
    if argv and argv[0] == "report" and len(argv) == 7:
        print(report(argv[1], int(argv[2]), float(argv[3]), float(argv[4]), argv[5], argv[6]))
        return 0
    print(__doc__, file=sys.stderr)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
