"""Summarise rocprofv3 CSV output of bench.py runs into a short markdown table and a JSON
file of HBM traffic (and the SQ counters, when given) per kernel launch.

Inputs are the output directories of up to four runs of the same command, each in CSV
(--output-format csv) or rocprofv3's default rocpd database:

  --trace DIR   rocprofv3 --kernel-trace --stats           (run_kernel_stats.csv)
  --fetch DIR   rocprofv3 --pmc FETCH_SIZE                 (run_counter_collection.csv)
  --write DIR   rocprofv3 --pmc WRITE_SIZE
  --sq DIR      rocprofv3 --pmc SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_BUSY_CYCLES ...; may be given
                more than once (e.g. a second pass with TA_BUSY_avr GRBM_GUI_ACTIVE)

FETCH_SIZE and WRITE_SIZE are KiB per dispatch.  Each is scaled by what the same counter
reads for the runtime's copy kernel (__amd_rocclr_copyBuffer) on copies of known size: the
model's copies are whole float4 buffers (16 B x max_n: positions, velocities, predicted
positions) or key buffers (4 B x max_n).  The copy dispatches are split at the largest gap
of their counter values into those two size classes, and the scale is the counter total
over the byte total.

    python -m realtimeparticles_amd.profile_summary --trace D1 --fetch D2 --write D3 \\
        --max-n 1049088 --command "..." --md out.md --json out.json
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import re
from collections import defaultdict

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def short_name(name: str) -> str:
    """'void rtp::(anonymous namespace)::k_correct<2>(float4 const*, ...)' -> 'k_correct<2>'."""
    s = name.strip('"')
    depth, cut = 0, len(s)
    for i, ch in enumerate(s):  # first '(' outside template brackets, after the name
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0 and i > 0 and s[i - 1] != ":":
            cut = i
            break
    s = s[:cut]
    s = re.sub(r"^void\s+", "", s)
    s = s.replace("(anonymous namespace)::", "")
    s = re.sub(r"^rtp::", "", s)
    return s


def base_name(name: str) -> str:
    return short_name(name).split("<")[0]


def read_csv(path: str):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def find(dirpath: str, suffix: str) -> str:
    for root, _, files in os.walk(dirpath):
        for fn in files:
            if fn.endswith(suffix):
                return os.path.join(root, fn)
    raise FileNotFoundError(f"no *{suffix} under {dirpath}")


def find_db(dirpath: str):
    """The rocpd database (rocprofv3's default output) under dirpath, or None."""
    for root, _, files in os.walk(dirpath):
        for fn in files:
            if fn.endswith("_results.db"):
                return os.path.join(root, fn)
    return None


def _db_rows(path: str, sql: str):
    """Rows of `sql` over the library's own process: a profiled command may start child
    processes (the bench's copy-bandwidth probe), whose dispatches are left out."""
    import sqlite3
    con = sqlite3.connect(path)
    try:
        cur = con.cursor()
        top = cur.execute("SELECT pid, COUNT(*) FROM kernels WHERE name LIKE '%rtp::%' "
                          "GROUP BY pid ORDER BY 2 DESC").fetchone()
        if top is None:
            return []
        return cur.execute(sql, (top[0],)).fetchall()
    finally:
        con.close()


def kernel_stats(trace_dir: str):
    db = None
    try:
        rows = read_csv(find(trace_dir, "kernel_stats.csv"))
    except FileNotFoundError:
        db = find_db(trace_dir)
        if db is None:
            raise
        res = _db_rows(db, "SELECT name, COUNT(*), SUM(end - start), AVG(end - start), "
                           "MIN(end - start), MAX(end - start) FROM kernels WHERE pid = ? "
                           "GROUP BY name")
        total = sum(r[2] for r in res) or 1.0
        rows = [{"Name": r[0], "Calls": r[1], "TotalDurationNs": r[2], "AverageNs": r[3],
                 "MinNs": r[4], "MaxNs": r[5], "Percentage": 100.0 * r[2] / total} for r in res]
    out = {}
    for r in rows:
        out[short_name(r["Name"])] = {"calls": int(r["Calls"]), "total_ns": float(r["TotalDurationNs"]),
                                      "avg_ns": float(r["AverageNs"]), "min_ns": float(r["MinNs"]),
                                      "max_ns": float(r["MaxNs"]), "percent": float(r["Percentage"])}
    return out


def counter_rows(pmc_dir: str):
    """Per-dispatch counter rows (CSV field names) of a --pmc run, CSV or rocpd output."""
    try:
        return read_csv(find(pmc_dir, "counter_collection.csv"))
    except FileNotFoundError:
        db = find_db(pmc_dir)
        if db is None:
            raise
    res = _db_rows(db, "SELECT kernel_name, counter_name, value, vgpr_count, sgpr_count, "
                       "lds_block_size, workgroup_size FROM counters_collection WHERE pid = ?")
    return [{"Kernel_Name": r[0], "Counter_Name": r[1], "Counter_Value": r[2], "VGPR_Count": r[3],
             "SGPR_Count": r[4], "LDS_Block_Size": r[5], "Workgroup_Size": r[6]} for r in res]


def counters(pmc_dir: str, counter: str):
    """{short kernel name: [value per dispatch]} and the dispatch metadata."""
    rows = counter_rows(pmc_dir)
    vals = defaultdict(list)
    meta = {}
    for r in rows:
        if r["Counter_Name"] != counter:
            continue
        k = short_name(r["Kernel_Name"])
        vals[k].append(float(r["Counter_Value"]))
        meta[k] = {"vgpr": int(r["VGPR_Count"]), "sgpr": int(r["SGPR_Count"]),
                   "lds": int(r["LDS_Block_Size"]), "workgroup": int(r["Workgroup_Size"])}
    return vals, meta


def copy_factor(vals, max_n: int) -> float | None:
    """Counter bytes per real byte of the copy kernel, from its whole-buffer copies."""
    import math
    v = [x for x in vals.get(COPY_KERNEL, []) if x > 0]
    if not v:
        return None
    top = max(v)
    v = sorted(x for x in v if x >= top / 8.0)  # whole-buffer copies only
    if len(v) < 2:
        return None
    logs = [math.log(x) for x in v]
    gaps = [logs[i + 1] - logs[i] for i in range(len(v) - 1)]
    cut = max(range(len(gaps)), key=gaps.__getitem__) + 1
    small, large = v[:cut], v[cut:]
    # the two classes must be the 4-byte and 16-byte copies: their counters differ by ~4x
    ratio = (sum(large) / len(large)) / (sum(small) / len(small))
    if not 3.0 < ratio < 5.0:
        return None
    real = (len(small) * 4.0 + len(large) * 16.0) * max_n
    return (sum(v) * 1024.0) / real


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--sq", action="append", default=[])
    ap.add_argument("--max-n", type=int, required=True)
    ap.add_argument("--command", default="")
    ap.add_argument("--title", default="rocprofv3 summary")
    ap.add_argument("--md", required=True)
    ap.add_argument("--json")
    a = ap.parse_args(argv)

    stats = kernel_stats(a.trace) if a.trace else {}
    fetch, meta = counters(a.fetch, "FETCH_SIZE") if a.fetch else ({}, {})
    write, meta_w = counters(a.write, "WRITE_SIZE") if a.write else ({}, {})
    meta = meta or meta_w
    ff = copy_factor(fetch, a.max_n) if fetch else None
    fw = copy_factor(write, a.max_n) if write else None
    sq = defaultdict(dict)
    for sqdir in a.sq:
        rows = counter_rows(sqdir)
        acc = defaultdict(lambda: defaultdict(float))
        cnt = defaultdict(lambda: defaultdict(int))
        for r in rows:
            k = short_name(r["Kernel_Name"])
            acc[k][r["Counter_Name"]] += float(r["Counter_Value"])
            cnt[k][r["Counter_Name"]] += 1
        for k, d in acc.items():
            sq[k].update({c: d[c] / cnt[k][c] for c in d})

    names = sorted(set(stats) | set(fetch) | set(write) | set(sq),
                   key=lambda k: -stats.get(k, {}).get("total_ns", 0.0))
    traffic = {}
    lines = [f"# {a.title}", "", "Commands (run from the repository root on one MI355X; "
             "`cd /tmp && export TMPDIR=/tmp` first for rocprofv3):", "", "```", a.command.strip(), "```", ""]
    if ff is not None or fw is not None:
        lines += [f"Copy-kernel scale (counter bytes per real byte of {COPY_KERNEL} on the model's "
                  f"whole-buffer copies, 16 B and 4 B x {a.max_n}): FETCH_SIZE "
                  f"{ff if ff is None else round(ff, 3)}, WRITE_SIZE {fw if fw is None else round(fw, 3)}. "
                  "Traffic columns are the counters divided by these scales.", ""]
    lines += ["| kernel | calls | avg us | total ms | % | FETCH MB/launch (scaled) | WRITE MB/launch (scaled) | VGPR | LDS B |",
              "|---|---|---|---|---|---|---|---|---|"]
    for k in names:
        st = stats.get(k, {})
        fv, wv = fetch.get(k), write.get(k)
        fb = (sum(fv) / len(fv) * 1024.0 / ff) if fv and ff else None
        wb = (sum(wv) / len(wv) * 1024.0 / fw) if wv and fw else None
        m = meta.get(k, {})
        if fb is not None or wb is not None or k in sq:
            traffic[k] = {"fetch_bytes": fb, "write_bytes": wb,
                          "traffic_bytes": (fb or 0.0) + (wb or 0.0), "base": base_name(k),
                          "avg_ns": st.get("avg_ns"), "calls": st.get("calls")}
            if k in sq:
                traffic[k]["counters"] = dict(sq[k])
        lines.append("| {} | {} | {} | {} | {} | {} | {} | {} | {} |".format(
            k, st.get("calls", ""), "" if not st else f"{st['avg_ns'] / 1e3:.2f}",
            "" if not st else f"{st['total_ns'] / 1e6:.3f}", "" if not st else f"{st['percent']:.1f}",
            "" if fb is None else f"{fb / 1e6:.2f}", "" if wb is None else f"{wb / 1e6:.2f}",
            m.get("vgpr", ""), m.get("lds", "")))
    if sq:
        cols = sorted({c for d in sq.values() for c in d})
        lines += ["", "Counters (mean per dispatch, summed over the GPU's instances):", "",
                  "| kernel | " + " | ".join(cols) + " |", "|---|" + "---|" * len(cols)]
        for k in names:
            if k in sq:
                lines.append(f"| {k} | " + " | ".join(
                    f"{sq[k][c]:.4g}" if c in sq[k] else "" for c in cols) + " |")
    os.makedirs(os.path.dirname(os.path.abspath(a.md)), exist_ok=True)
    with open(a.md, "w") as f:
        f.write("\n".join(lines) + "\n")
    if a.json:
        with open(a.json, "w") as f:
            json.dump({"command": a.command.strip(), "max_n": a.max_n, "fetch_scale": ff,
                       "write_scale": fw, "kernels": traffic}, f, indent=1)


if __name__ == "__main__":
    main()
