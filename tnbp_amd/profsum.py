"""Summaries of rocprofv3 output as short markdown (kernel-trace statistics
and per-dispatch PMC counters), so that only a few KB of text need keeping.

    python -m tnbp_amd.profsum stats <results.db> --title T --cmd "C"
    python -m tnbp_amd.profsum pmc <counter_collection.csv> --title T --cmd "C"
    python -m tnbp_amd.profsum util <counter_collection.csv> --title T --cmd "C"
    python -m tnbp_amd.profsum traffic <fetch.csv> <write.csv> <probe_fetch.csv> \
        <probe_write.csv> --pattern REGEX --precision P --key K --alg-bytes B --launch TEXT
"""
from __future__ import annotations

import argparse
import csv
import json
import re
import sqlite3
from collections import defaultdict


def short(name):
    n = re.sub(r"\(anonymous namespace\)::", "", name)
    n = re.sub(r"HIP_vector_type<(\w+), 2u>", r"\g<1>2", n)
    n = re.sub(r"hbp::Cplx<(\w+)>::type", r"C<\1>", n)
    m = re.match(r"(?:void )?([\w:]+(?:<[^()]*>)?)\(", n)
    return m.group(1) if m else n[:90]


def stats(db, title, cmd, top=25):
    con = sqlite3.connect(db)
    rows = con.execute("select name, total_calls, total_duration, average, percentage "
                       "from top_kernels").fetchall()
    out = [f"# {title}\n", f"Command (repository root):\n\n```\n{cmd}\n```\n"]
    # the view's durations are in microseconds (its averages match HIP event
    # timings of the same launches)
    tot = sum(r[2] for r in rows)
    out.append(f"Total kernel time {tot / 1e6:.3f} s over {sum(r[1] for r in rows)} dispatches.\n")
    out.append("| kernel | calls | total ms | avg us | % |\n|---|---|---|---|---|")
    for n, c, t, a, p in rows[:top]:
        out.append(f"| `{short(n)}` | {c} | {t / 1e3:.1f} | {a:.1f} | {p:.1f} |")
    return "\n".join(out) + "\n"


def pmc(path, title, cmd):
    """Mean and total of every counter per kernel (FETCH_SIZE / WRITE_SIZE are
    in KiB per dispatch)."""
    agg = defaultdict(lambda: [0, 0.0, 0.0, float("inf"), 0.0])
    with open(path) as f:
        for r in csv.DictReader(f):
            key = (short(r["Kernel_Name"]), r["Counter_Name"])
            v = float(r["Counter_Value"])
            a = agg[key]
            a[0] += 1
            a[1] += v
            a[3] = min(a[3], v)
            a[4] = max(a[4], v)
    out = [f"# {title}\n", f"Command (repository root):\n\n```\n{cmd}\n```\n",
           "| kernel | counter | dispatches | mean | min | max |\n|---|---|---|---|---|---|"]
    for (k, c), (n, s, _, lo, hi) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        out.append(f"| `{k}` | {c} | {n} | {s / n:.1f} | {lo:.1f} | {hi:.1f} |")
    return "\n".join(out) + "\n"


def util(path, title, cmd, simds=1024, xcds=8, min_us=20.0):
    """Per kernel and grid size (dispatches longer than min_us): mean duration, the shader
    clock (GRBM_GUI_ACTIVE, which rocprofv3 sums over the XCDs, / XCDs /
    duration), the fraction of SIMD cycles the MFMA unit is busy
    (SQ_VALU_MFMA_BUSY_CYCLES / (GUI_ACTIVE per XCD x SIMDs)), the MFMA flops
    (SQ_INSTS_VALU_MFMA_MOPS_* x 512) and LDS bank-conflict cycles per active
    LDS cycle."""
    disp = defaultdict(dict)
    meta = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            d = int(r["Dispatch_Id"])
            disp[d][r["Counter_Name"]] = float(r["Counter_Value"])
            meta[d] = ((short(r["Kernel_Name"]), int(r["Grid_Size"])),
                       (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9)
    agg = defaultdict(lambda: defaultdict(float))
    for d, c in disp.items():
        name, dur = meta[d]
        if dur * 1e6 < min_us:
            continue
        a = agg[name]
        a["n"] += 1
        a["dur"] += dur
        for k, v in c.items():
            a[k] += v
    out = [f"# {title}\n", f"Command (repository root):\n\n```\n{cmd}\n```\n",
           f"Dispatches of at least {min_us:g} us; {xcds} XCDs, {simds} SIMDs.\n",
           "| kernel | grid | dispatches | mean us | clock GHz | MFMA busy | MFMA GFLOP / "
           "dispatch | MFMA TFLOP/s | LDS conflict / active |\n|---|---|---|---|---|---|---|---|---|"]
    for (name, grid), a in sorted(agg.items(), key=lambda kv: -kv[1]["dur"]):
        n, dur = a["n"], a["dur"] / a["n"]
        gui = a.get("GRBM_GUI_ACTIVE", 0.0) / n / xcds
        busy = a.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0) / n
        mops = sum(v for k, v in a.items() if k.startswith("SQ_INSTS_VALU_MFMA_MOPS")) / n * 512
        lds = a.get("SQ_LDS_IDX_ACTIVE", 0.0)
        conf = a.get("SQ_LDS_BANK_CONFLICT", 0.0) / lds if lds else float("nan")
        out.append(f"| `{name}` | {grid} | {int(n)} | {dur * 1e6:.1f} | {gui / dur / 1e9:.3f} | "
                   f"{busy / (gui * simds) if gui else float('nan'):.3f} | {mops / 1e9:.2f} | "
                   f"{mops / dur / 1e12:.1f} | {conf:.3f} |")
    return "\n".join(out) + "\n"


def _max_counter(path, pattern):
    best = 0.0
    with open(path) as f:
        for r in csv.DictReader(f):
            if re.search(pattern, short(r["Kernel_Name"])):
                best = max(best, float(r["Counter_Value"]))
    return best


def traffic(fetch, write, probe_fetch, probe_write, pattern, precision, key, alg_bytes, launch,
            cmds, out_path):
    """Merge the HBM bytes of the largest dispatch matching `pattern` into the
    JSON at out_path: FETCH_SIZE and WRITE_SIZE (KiB) each scaled by the
    factor that maps the counter to the known bytes of the streaming probe."""
    pf, pw = _max_counter(probe_fetch, "k_scale"), _max_counter(probe_write, "k_scale")
    # the probe reads and writes 512 MiB per dispatch (python -m tnbp_amd.traffic_probe)
    known_kib = 512 * 1024
    ff, wf = known_kib / pf, known_kib / pw
    fk, wk = _max_counter(fetch, pattern), _max_counter(write, pattern)
    try:
        with open(out_path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        table = {}
    table.setdefault("calibration", {})[precision] = {
        "probe": "k_scale<double2,double>: 536870912 B read + 536870912 B written per dispatch",
        "FETCH_SIZE_KiB": pf, "WRITE_SIZE_KiB": pw, "fetch_factor": ff, "write_factor": wf}
    table.setdefault(precision, {})[key] = {
        "launch": launch, "fetch_KiB": fk, "write_KiB": wk,
        "hbm_bytes_per_launch": round(fk * 1024 * ff + wk * 1024 * wf),
        "algorithmic_bytes": alg_bytes, "commands": cmds}
    with open(out_path, "w") as f:
        json.dump(table, f, indent=1)
    return table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("kind", choices=["stats", "pmc", "util", "traffic"])
    ap.add_argument("path", nargs="+")
    ap.add_argument("--title", default="rocprofv3 summary")
    ap.add_argument("--cmd", default=None, action="append")
    ap.add_argument("--pattern", default="")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--key", default="")
    ap.add_argument("--alg-bytes", type=float, default=0.0)
    ap.add_argument("--launch", default="")
    ap.add_argument("--out", default="profiles/traffic.json")
    a = ap.parse_args()
    cmds = [c for c in (a.cmd or []) if c]
    cmd = "\n".join(cmds)
    if a.kind == "stats":
        print(stats(a.path[0], a.title, cmd), end="")
    elif a.kind == "pmc":
        print(pmc(a.path[0], a.title, cmd), end="")
    elif a.kind == "util":
        print(util(a.path[0], a.title, cmd), end="")
    else:
        t = traffic(*a.path[:4], a.pattern, a.precision, a.key, a.alg_bytes, a.launch,
                    cmds, a.out)
        print(json.dumps(t[a.precision][a.key]))


if __name__ == "__main__":
    main()
