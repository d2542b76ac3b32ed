"""Summaries of rocprofv3 output for the step's kernels.

``python -m dasha_amd.profiling --stats DIR [--fetch DIR --write DIR]
--workload NAME --command "..." --out profiles/NAME.md`` turns the CSV files
of a ``rocprofv3 --kernel-trace --stats`` run and of two ``--pmc
FETCH_SIZE`` / ``--pmc WRITE_SIZE`` runs into a short markdown summary and,
with the counters, ``profiles/traffic_NAME.json``: HBM bytes per launch of
each kernel family, scaled by what the same counters read for the library's
plain copy kernel ``k_copy`` (bytes known: the probe copies ``--copy-bytes``
per launch, read and write).  ``--sq DIR`` instead summarises one
``rocprofv3 --pmc`` run of eight SQ counters per wave of each step kernel.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import re

# kernel symbol (demangled prefix) -> family name used by bench.py
_FAMILIES = [
    (r"k_curl\b", "curl"),
    (r"k_zsmooth<2, 0, 0, -1[,>]", "avrg_radial_x2"),
    (r"k_sweepz<3, 0, 0, 3, 4, -1>", "avrg_radial_x1"),
    (r"k_zsmooth<2, 1, 0, -1[,>]", "avrgh_x2"),
    (r"k_sweepz<3, 1, 0, 3, 4, -1>", "avrgh_x1"),
    (r"k_zsmooth<3, 1, 0, -2[,>]", "avrgh_x3_addE"),
    (r"k_zsmooth<2, 2, 1, -1[,>]", "avrg_state_x2"),
    (r"k_sweepz<8, 2, 1, 8, 4, -1>", "avrg_state_x1"),
    (r"k_zsmooth<2, 2, 1, [0-4][,>]", "avrg_state_x2_rk"),
    (r"k_sweepz<8, 2, 1, 8, 4, [0-3]>", "avrg_state_x1_rk"),
    (r"k_zsmooth<2, 2, 2, -1[,>]", "avrgtt_x2_eq"),
    (r"k_rk\b", "rk_update"),
    (r"k_rhs_a\b", "rhs_flux_E_hall"),
    (r"k_rhs_b<|k_rhs_b2<|k_fill_state\b", "rhs_curlE_pressure"),
    (r"k_smooth_tt<", "avrgtt"),
    (r"k_bc_tt\b", "bndrytt"),
    (r"k_add_eq\b", "post_add_eq"),
    (r"k_dt_(partial|final)\b", "setdt"),
    (r"k_copy\b", "copy_probe"),
    # magnetosphere mode (one sweep per launch, fills of two launches)
    (r"km_sweep<0>", "mag_avrg_radial"),
    (r"km_sweep<1>", "mag_avrgh"),
    (r"km_sweep<2>", "mag_avrg_state"),
    (r"km_fill_(x|y|z|yz)<|km_mask\b", "mag_fill"),
    (r"km_curl\b", "mag_curl"),
    (r"km_flux\b|km_rhs_a\b", "mag_rhs_flux_E"),
    (r"km_hall\b|km_add3\b", "mag_hall"),
    (r"km_rhs_b\b", "mag_rhs_curlE_pressure"),
    (r"km_eq\b", "mag_post_eq"),
    # discrete adjoint (reverse sweep)
    (r"k_adj_filter<", "adj_sweep"),
    (r"k_adj_bc_[xz]\b", "adj_fill"),
    (r"k_adj_curle_p\b", "adj_curlE_pressure"),
    (r"k_adj_hall\b", "adj_hall"),
    (r"k_adj_ext_flux\b", "adj_ext_flux"),
    (r"k_adj_curl\b", "adj_curl"),
    (r"k_adj_rk\b", "adj_rk"),
    (r"k_adj_objective\b", "adj_objective"),
]


# families whose kernels are launched together for one call of the family
_TOGETHER = {"rhs_curlE_pressure", "setdt", "mag_fill", "mag_rhs_flux_E"}


def launch_group(kernel_name: str) -> str:
    """kernels of one family launched together share a group: the smoothing
    sweeps run as two launches, the windows without face rules (launch-mode
    template argument 1) and those with them (2, optionally followed by the
    face launch's row count)"""
    inst = re.sub(r"\(.*", "", kernel_name.replace("(anonymous namespace)::", ""))
    return re.sub(r", [12](, \d+)?>$", ">", inst)


def family(kernel_name: str) -> str | None:
    for pat, fam in _FAMILIES:
        if re.search(pat, kernel_name):
            return fam
    return None


def _one(d: str, pattern: str) -> str:
    hits = glob.glob(os.path.join(d, "**", pattern), recursive=True)
    if not hits:
        raise FileNotFoundError(f"no {pattern} under {d}")
    return hits[0]


def kernel_stats(d: str):
    rows = list(csv.DictReader(open(_one(d, "*kernel_stats.csv"))))
    out = []
    for r in rows:
        out.append(dict(name=r["Name"], family=family(r["Name"]), calls=int(r["Calls"]),
                        avg_us=float(r["AverageNs"]) / 1e3, total_ms=float(r["TotalDurationNs"]) / 1e6,
                        pct=float(r["Percentage"])))
    return out


def family_calls(d: str):
    """wall time of each call of each kernel family from the kernel trace:
    a family's dispatches in start order, a new call where a kernel of the
    current call would repeat, an alternative of the family (another RK
    stage) begins or more than 20 us pass after the call's last end (the
    kernels of one call -- a sweep's face and
    face-free launches, the parts of right part 2 -- may run side by side on
    two streams); wall = last end - first start"""
    hits = glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True)
    if not hits:
        return {}
    disp = collections.defaultdict(list)
    for r in csv.DictReader(open(hits[0])):
        fam = family(r["Kernel_Name"])
        if fam is None or fam == "copy_probe":
            continue
        disp[fam].append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    out = {}
    for fam, ds in disp.items():
        ds.sort()
        calls, cur, names, grp = [], None, set(), None
        for s, e, n in ds:
            g = "*" if fam in _TOGETHER else launch_group(n)
            # a kernel of one call starts before, or within 20 us of, the end
            # of the call's other kernels
            if cur is None or n in names or g != grp or s > cur[1] + 20000:
                if cur is not None:
                    calls.append(cur[1] - cur[0])
                cur, names, grp = [s, e], set(), g
            cur[0], cur[1] = min(cur[0], s), max(cur[1], e)
            names.add(n)
        if cur is not None:
            calls.append(cur[1] - cur[0])
        out[fam] = (len(calls), sum(calls) / len(calls) / 1e6)
    return out


def counter_per_kernel(d: str, counter: str):
    """mean value of `counter` per dispatch, by kernel name (KiB)"""
    acc = collections.defaultdict(list)
    for r in csv.DictReader(open(_one(d, "*counter_collection.csv"))):
        if r["Counter_Name"] == counter:
            acc[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def traffic(fetch_dir: str, write_dir: str, copy_bytes: int):
    """bytes per launch by family: FETCH_SIZE and WRITE_SIZE (KiB) scaled so
    that the copy kernel reads and writes copy_bytes"""
    fetch = counter_per_kernel(fetch_dir, "FETCH_SIZE")
    write = counter_per_kernel(write_dir, "WRITE_SIZE")
    ck = [k for k in fetch if family(k) == "copy_probe"]
    if not ck:
        raise RuntimeError("the copy kernel is missing from the counter runs")
    f_scale = copy_bytes / (fetch[ck[0]] * 1024.0)
    w_scale = copy_bytes / (write[ck[0]] * 1024.0)
    fams = collections.defaultdict(lambda: collections.defaultdict(lambda: [0.0, 0.0]))
    for k in fetch:
        fam = family(k)
        if fam is None or fam == "copy_probe" or k not in write:
            continue
        a = fams[fam]["*" if fam in _TOGETHER else launch_group(k)]
        a[0] += fetch[k] * 1024.0 * f_scale
        a[1] += write[k] * 1024.0 * w_scale
    # kernels launched together inside one timed bracket add up; the
    # family's alternatives (RK stages) average
    res = {}
    for fam, groups in fams.items():
        n = len(groups)
        rd = sum(v[0] for v in groups.values()) / n
        wr = sum(v[1] for v in groups.values()) / n
        res[fam] = {"read_bytes_per_launch": rd, "write_bytes_per_launch": wr, "bytes_per_launch": rd + wr}
    return res, f_scale, w_scale


_SQ = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_INSTS_SALU", "SQ_WAVE_CYCLES", "SQ_WAIT_INST_ANY",
       "SQ_LDS_BANK_CONFLICT", "SQ_BUSY_CYCLES")


def sq_table(d: str, pattern: str = r"k_zsmooth|k_rhs_a\b|k_curl\b|k_rhs_b2"):
    """per kernel (matching pattern): the SQ counters of one --pmc run,
    averaged over the kernel's dispatches, per wave (SQ_WAVES per dispatch)"""
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(_one(d, "*counter_collection.csv"))):
        if re.search(pattern, r["Kernel_Name"]):
            acc[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    rows = []
    for k, cs in acc.items():
        mean = {c: sum(v) / len(v) for c, v in cs.items()}
        w = mean.get("SQ_WAVES", 0.0)
        if w <= 0:
            continue
        nm = re.sub(r"\(.*", "", k.replace("(anonymous namespace)::", ""))
        rows.append((nm, w, {c: mean.get(c, 0.0) / w for c in _SQ[1:]}))
    return sorted(rows, key=lambda r: r[0])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats")
    ap.add_argument("--sq", help="directory of a rocprofv3 --pmc run of the SQ counters")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-bytes", type=int, default=2 << 30)
    ap.add_argument("--workload", required=True)
    ap.add_argument("--command", required=True, help="the rocprofv3 command lines, verbatim")
    ap.add_argument("--out", required=True)
    ap.add_argument("--cells", type=int, default=0, help="physical cells (for bytes per cell)")
    args = ap.parse_args()
    if args.sq:
        lines = [f"# {args.workload}: SQ counters (rocprofv3)", "", "Command:", "", "```"]
        lines += args.command.split("\\n")
        lines += ["```", "", "Template arguments of k_zsmooth: <passes, coefficient kind, fill type, RK stage, "
                  "launch class (0 all windows, 1 face-free, 2 face), rows per block>.  Per dispatch, averaged "
                  "over the kernel's dispatches, divided by SQ_WAVES; SQ_WAVE_CYCLES counts 4-cycle units.", "",
                  "| kernel | waves per launch | VALU / wave | LDS / wave | SALU / wave | wave cycles / wave "
                  "| waiting / wave | LDS bank conflicts / wave | busy cycles / wave |",
                  "|---|---|---|---|---|---|---|---|---|"]
        for nm, w, v in sq_table(args.sq):
            lines.append(f"| `{nm}` | {w:.0f} | {v['SQ_INSTS_VALU']:.0f} | {v['SQ_INSTS_LDS']:.0f} | "
                         f"{v['SQ_INSTS_SALU']:.0f} | {v['SQ_WAVE_CYCLES']:.0f} | {v['SQ_WAIT_INST_ANY']:.0f} | "
                         f"{v['SQ_LDS_BANK_CONFLICT']:.1f} | {v['SQ_BUSY_CYCLES']:.1f} |")
        with open(args.out, "w") as f:
            f.write("\n".join(lines) + "\n")
        print(args.out)
        return
    if not args.stats:
        ap.error("--stats or --sq is required")
    ks = kernel_stats(args.stats)
    lines = [f"# {args.workload}: kernel times (rocprofv3)", "", "Commands:", "", "```"]
    lines += args.command.split("\\n")
    lines += ["```", "", "| kernel family | kernel | calls | avg us | total ms | % |", "|---|---|---|---|---|---|"]
    for r in sorted(ks, key=lambda r: -r["total_ms"]):
        nm = re.sub(r"\(.*", "", r["name"].replace("(anonymous namespace)::", ""))
        lines.append(f"| {r['family'] or '-'} | `{nm}` | {r['calls']} | {r['avg_us']:.1f} | "
                     f"{r['total_ms']:.3f} | {r['pct']:.1f} |")
    # per family: kernels launched together add up, alternatives count apart
    fam_t = collections.defaultdict(lambda: collections.defaultdict(lambda: [0.0, 0]))
    for r in ks:
        if r["family"] is None or r["family"] == "copy_probe":
            continue
        g = fam_t[r["family"]]["*" if r["family"] in _TOGETHER else launch_group(r["name"])]
        g[0] += r["total_ms"]
        g[1] = max(g[1], r["calls"])
    walls = family_calls(args.stats)
    lines += ["", "Per call of each kernel family (the kernels one call of the family runs; "
              "what bench.py brackets with HIP events). Summed: the kernels' durations added; "
              "wall: first start to last end of the call in the trace (the face and face-free "
              "launches of a sweep and the two parts of right part 2 run side by side on two "
              "streams; the first two Hall passes run beside the smoothing of E):", "",
              "| kernel family | calls | summed avg ms | wall avg ms |", "|---|---|---|---|"]
    for fam, groups in sorted(fam_t.items(), key=lambda kv: -sum(v[0] for v in kv[1].values())):
        calls = sum(v[1] for v in groups.values())
        w = walls.get(fam)
        lines.append(f"| {fam} | {calls} | {sum(v[0] for v in groups.values()) / max(calls, 1):.3f} | "
                     f"{'%.3f' % w[1] if w else '-'} |")
    if args.fetch and args.write:
        tr, fs, ws = traffic(args.fetch, args.write, args.copy_bytes)
        lines += ["", "HBM traffic per launch (FETCH_SIZE x %.3f, WRITE_SIZE x %.3f: the factors that make "
                  "the copy kernel read and write %d bytes)" % (fs, ws, args.copy_bytes), "",
                  "| family | read MB | write MB | total MB |" + (" B/cell |" if args.cells else ""),
                  "|---|---|---|---|" + ("---|" if args.cells else "")]
        for fam, v in sorted(tr.items()):
            row = (f"| {fam} | {v['read_bytes_per_launch'] / 1e6:.1f} | {v['write_bytes_per_launch'] / 1e6:.1f} | "
                   f"{v['bytes_per_launch'] / 1e6:.1f} |")
            if args.cells:
                row += f" {v['bytes_per_launch'] / args.cells:.0f} |"
            lines.append(row)
        jpath = os.path.join(os.path.dirname(args.out), f"traffic_{args.workload}.json")
        with open(jpath, "w") as f:
            json.dump({"workload": args.workload, "fetch_scale": fs, "write_scale": ws,
                       "kernels": tr}, f, indent=1, sort_keys=True)
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(args.out)


if __name__ == "__main__":
    main()
