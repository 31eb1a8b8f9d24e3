"""Summaries of rocprofv3 runs for profiles/.

traffic: HBM bytes per kernel launch from two counter runs (FETCH_SIZE, WRITE_SIZE).  The
counters are reported in KiB per dispatch; each is scaled by what the same counter reads
for a device-to-device copy of known size (256 MiB, issued by ``bench.py --profile-only``),
so the result is in bytes as a plain copy would count them.

  python -m ffdas_amd.utils.pmc traffic <fetch_run_dir> <write_run_dir> --out profiles/<name>

writes ``<name>.json`` (per-kernel bytes, used by bench.py for ``roofline.traffic``) and
``<name>.md``.

stats: the kernel table of a ``rocprofv3 --kernel-trace --stats --output-format csv`` run.

counters: per-kernel utilisation (VALU, LDS bank conflicts, L1/L2 hit rates, TA) from one
``rocprofv3 --pmc`` run per pass of COUNTER_PASSES, each under <base>/<pass>/.

  python -m ffdas_amd.utils.pmc stats <run_dir> --out profiles/<name> --command "..."
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import re
import statistics

COPY_BYTES = 256 * 1024 * 1024


def short_name(kernel: str) -> str:
    m = re.search(r"ffdas::(?:\(anonymous namespace\)::)?(?:dev::)?(\w+)", kernel)
    if m:
        return m.group(1)
    m = re.search(r"_ZN5ffdas3dev\d+([A-Za-z_]\w*?)(?:I|E)", kernel)
    if m:
        return m.group(1)
    if "copyBuffer" in kernel:
        return "copyBuffer"
    return kernel.split("(")[0][-60:]


def read_counter(run_dir: str, counter: str, variants: bool = False) -> dict:
    """{short kernel name: [values per dispatch]} for one counter; with variants,
    {short kernel name: {full kernel name: [values per dispatch]}} (template instances apart)."""
    files = glob.glob(os.path.join(run_dir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {run_dir}")
    per = collections.defaultdict(float)
    names = {}
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] != counter:
                    continue
                key = (f, int(r["Dispatch_Id"]))
                per[key] += float(r["Counter_Value"])
                names[key] = (short_name(r["Kernel_Name"]), r["Kernel_Name"])
    if variants:
        nested = collections.defaultdict(lambda: collections.defaultdict(list))
        for key in sorted(per):
            nested[names[key][0]][names[key][1]].append(per[key])
        return nested
    out = collections.defaultdict(list)
    for key in sorted(per):
        out[names[key][0]].append(per[key])
    return out


def summarize(fetch_dir: str, write_dir: str) -> dict:
    fetch = read_counter(fetch_dir, "FETCH_SIZE", variants=True)
    write = read_counter(write_dir, "WRITE_SIZE", variants=True)
    if "copyBuffer" not in fetch or "copyBuffer" not in write:
        raise RuntimeError("the calibration copy (copyBuffer) is missing from the runs")
    f_copy = max(max(v) for v in fetch["copyBuffer"].values()) * 1024.0
    w_copy = max(max(v) for v in write["copyBuffer"].values()) * 1024.0
    kf, kw = COPY_BYTES / f_copy, COPY_BYTES / w_copy

    def per_launch(runs):
        # the median dispatch of each template instance (warm-ups and odd-sized calls
        # aside), summed over the instances: one operation that launches a kernel in two
        # instances (grad_x's whole rounds and its half-frame last round) counts both
        return sum(statistics.median(v) for v in runs.values()) * 1024.0

    kernels = {}
    for name in sorted(set(fetch) | set(write)):
        if name == "copyBuffer" or not name.endswith("kernel"):
            continue
        fb = per_launch(fetch[name]) * kf if name in fetch else None
        wb = per_launch(write[name]) * kw if name in write else None
        kernels[name] = {"fetch_bytes": fb, "write_bytes": wb,
                         "traffic_bytes": (fb or 0.0) + (wb or 0.0),
                         "dispatches": sum(len(v) for v in fetch.get(name, {}).values()),
                         "instances": len(fetch.get(name, {}))}
    return {"calibration": {"copy_bytes": COPY_BYTES, "fetch_counter_bytes": f_copy,
                            "write_counter_bytes": w_copy, "fetch_factor": kf,
                            "write_factor": kw},
            "kernels": kernels}


def to_markdown(s: dict, command: str) -> str:
    c = s["calibration"]
    lines = ["# HBM traffic per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE)", "",
             f"Command: `{command}`", "",
             f"Calibration: a {c['copy_bytes'] / 2**20:.0f} MiB device copy reads "
             f"FETCH_SIZE = {c['fetch_counter_bytes'] / 2**20:.1f} MiB and WRITE_SIZE = "
             f"{c['write_counter_bytes'] / 2**20:.1f} MiB; factors {c['fetch_factor']:.3f} "
             f"(fetch) and {c['write_factor']:.3f} (write) are applied below.", "",
             "| kernel | fetch MB | write MB | traffic MB |", "|---|---|---|---|"]
    for k, v in s["kernels"].items():
        fb = v["fetch_bytes"] or 0.0
        wb = v["write_bytes"] or 0.0
        n = v.get("instances", 1)
        name = k if n <= 1 else f"{k} ({n} template instances per operation, summed)"
        lines.append(f"| {name} | {fb / 1e6:.1f} | {wb / 1e6:.1f} | {(fb + wb) / 1e6:.1f} |")
    return "\n".join(lines) + "\n"


def kernel_stats(run_dir: str, top: int = 20) -> str:
    """Markdown table of *kernel_stats.csv (calls, average and total time, share)."""
    files = glob.glob(os.path.join(run_dir, "**", "*kernel_stats.csv"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no kernel_stats.csv under {run_dir}")
    with open(files[0]) as fh:
        rows = list(csv.DictReader(fh))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = ["| kernel | calls | avg us | min us | max us | share % |", "|---|---|---|---|---|---|"]
    for r in rows[:top]:
        name = r["Name"]
        if "ffdas::" in name:
            name = name.replace("(anonymous namespace)::", "").split("(")[0]
        else:
            name = name[:90]
        out.append(f"| `{name}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                   f"{float(r['Percentage']):.2f} |")
    return "\n".join(out) + "\n"


# counter passes of `counters` (at most 8 SQ, 4 TCC, 4 TCP, 2 TA counters per pass)
COUNTER_PASSES = {
    "sq1": "SQ_WAVES SQ_INSTS_VALU SQ_INSTS_LDS SQ_BUSY_CYCLES SQ_WAVE_CYCLES SQ_INSTS_SALU "
           "SQ_INSTS_VMEM_RD GRBM_GUI_ACTIVE",
    "sq2": "SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_LDS SQ_LDS_BANK_CONFLICT SQ_WAIT_INST_LDS",
    "cache": "TCC_HIT_sum TCC_MISS_sum TCP_TOTAL_CACHE_ACCESSES_sum TCP_TCC_READ_REQ_sum "
             "TA_BUSY_avr",
}


def counter_values(base: str, cus: int = 256, simds: int = 4, xcds: int = 8) -> dict:
    """Per-kernel utilisation from the passes of COUNTER_PASSES under base/<pass>/ (the last
    dispatch of each kernel; passes that were not run are skipped): duration from
    GRBM_GUI_ACTIVE (summed over the XCDs), VALU busy = SQ_INSTS_VALU * 4 cycles / (CUs *
    SIMDs) / cycles, the share of cycles with an LDS bank conflict per CU, L2 (TCC) and L1
    (TCP) hit rates, TA busy."""
    vals = collections.defaultdict(dict)
    for name, counters in COUNTER_PASSES.items():
        if not os.path.isdir(os.path.join(base, name)):
            continue
        for c in counters.split():
            for kern, v in read_counter(os.path.join(base, name), c).items():
                vals[kern][c] = v[-1]
    out = {}
    nan = float("nan")
    for kern in sorted(vals):
        v = vals[kern]
        if "GRBM_GUI_ACTIVE" not in v or not kern.endswith("kernel"):
            continue
        cyc = v["GRBM_GUI_ACTIVE"] / xcds
        hit, miss = v.get("TCC_HIT_sum"), v.get("TCC_MISS_sum")
        acc = v.get("TCP_TOTAL_CACHE_ACCESSES_sum")
        out[kern] = {
            "us": cyc / 2.4e3,
            "valu_busy": v["SQ_INSTS_VALU"] * 4 / (cus * simds) / cyc if "SQ_INSTS_VALU" in v else nan,
            "lds_instr_per_cu": v["SQ_INSTS_LDS"] / cus if "SQ_INSTS_LDS" in v else nan,
            "lds_bank_conflict": v["SQ_LDS_BANK_CONFLICT"] / cus / cyc if "SQ_LDS_BANK_CONFLICT" in v else nan,
            "l2_hit": hit / (hit + miss) if hit is not None and miss is not None and hit + miss else nan,
            "l1_hit": 1.0 - v.get("TCP_TCC_READ_REQ_sum", 0.0) / acc if acc else nan,
            "ta_busy": v["TA_BUSY_avr"] / cyc if "TA_BUSY_avr" in v else nan,
        }
    return out


def counter_table(base: str, cus: int = 256, simds: int = 4, xcds: int = 8) -> str:
    """Markdown table of counter_values."""
    lines = ["| kernel | us | VALU busy | LDS instr/CU | LDS bank-conflict cycles | "
             "L2 hit | L1 hit | TA busy |", "|---|---|---|---|---|---|---|---|"]
    for kern, c in counter_values(base, cus, simds, xcds).items():
        lines.append(f"| {kern} | {c['us']:.0f} | {c['valu_busy']:.2f} | "
                     f"{c['lds_instr_per_cu']:.3g} | {c['lds_bank_conflict']:.2f} | "
                     f"{c['l2_hit']:.2f} | {c['l1_hit']:.2f} | {c['ta_busy']:.2f} |")
    return "\n".join(lines) + "\n"


# vector-memory passes of `vmem` (at most 4 TCP, 2 TA, 2 TD counters per pass)
VMEM_PASSES = {
    "l1a": "TCP_PENDING_STALL_CYCLES_sum TCP_READ_TAGCONFLICT_STALL_CYCLES_sum "
           "TCP_TCR_TCP_STALL_CYCLES_sum TCP_UTCL1_TRANSLATION_MISS_sum "
           "TA_ADDR_STALLED_BY_TC_CYCLES_sum TA_BUFFER_READ_WAVEFRONTS_sum GRBM_GUI_ACTIVE",
    "l1b": "TCP_TCC_READ_REQ_LATENCY_sum TCP_TCC_READ_REQ_sum TCP_TOTAL_CACHE_ACCESSES_sum "
           "TCP_TCP_TA_DATA_STALL_CYCLES_sum TA_TA_BUSY_sum TA_FLAT_READ_WAVEFRONTS_sum "
           "TD_TD_BUSY_sum TD_TC_STALL_sum GRBM_GUI_ACTIVE",
}


def vmem_table(base: str, cus: int = 256, xcds: int = 8) -> str:
    """Per-kernel vector-memory pipeline from the passes of VMEM_PASSES under base/<pass>/
    (the last dispatch of each kernel), per CU and cycle: vector L1 (TCP) cache accesses,
    load instructions (TA flat and buffer read wavefronts), accesses per instruction, texture-address
    (TA) and texture-data (TD) busy shares, their stalls on the L1, the L1's tag-conflict and
    pending-miss stalls, and the mean L1-to-L2 read latency in cycles."""
    vals = collections.defaultdict(dict)
    for name, counters in VMEM_PASSES.items():
        for c in counters.split():
            for kern, v in read_counter(os.path.join(base, name), c).items():
                vals[kern][c] = v[-1]
    lines = ["| kernel | us | L1 accesses/clk | loads/clk | accesses/load | TA busy | "
             "TA addr stalled by L1 | TD busy | TD stalled by L1 | L1 tag-conflict stall | "
             "L1 pending stall | L2 read latency (clk) |",
             "|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for kern in sorted(vals):
        v = vals[kern]
        if "GRBM_GUI_ACTIVE" not in v or not kern.endswith("kernel"):
            continue
        cyc = v["GRBM_GUI_ACTIVE"] / xcds
        per = lambda c: v.get(c, 0.0) / cus / cyc
        acc = per("TCP_TOTAL_CACHE_ACCESSES_sum")
        loads = per("TA_FLAT_READ_WAVEFRONTS_sum") + per("TA_BUFFER_READ_WAVEFRONTS_sum")
        req = v.get("TCP_TCC_READ_REQ_sum", 0.0)
        lat = v.get("TCP_TCC_READ_REQ_LATENCY_sum", 0.0) / req if req else float("nan")
        lines.append(
            f"| {kern} | {cyc / 2.4e3:.0f} | {acc:.2f} | {loads:.3f} | "
            f"{acc / loads if loads else float('nan'):.1f} | {per('TA_TA_BUSY_sum'):.2f} | "
            f"{per('TA_ADDR_STALLED_BY_TC_CYCLES_sum'):.2f} | {per('TD_TD_BUSY_sum'):.2f} | "
            f"{per('TD_TC_STALL_sum'):.2f} | {per('TCP_READ_TAGCONFLICT_STALL_CYCLES_sum'):.2f} | "
            f"{per('TCP_PENDING_STALL_CYCLES_sum'):.2f} | {lat:.0f} |")
    return "\n".join(lines) + "\n"


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--out", required=True, help="output path without extension")
    t.add_argument("--command", default="", help="the profiled command, for the summary")
    k = sub.add_parser("stats")
    k.add_argument("run_dir")
    k.add_argument("--out", required=True, help="output path without extension")
    k.add_argument("--command", default="", help="the profiled command, for the summary")
    k.add_argument("--title", default="rocprofv3 kernel statistics")
    c = sub.add_parser("counters")
    c.add_argument("base", help="directory holding one run per pass: sq1/ sq2/ cache/")
    c.add_argument("--out", required=True, help="output path without extension")
    c.add_argument("--command", default="", help="the profiled command, for the summary")
    m = sub.add_parser("vmem")
    m.add_argument("base", help="directory holding one run per pass: l1a/ l1b/")
    m.add_argument("--out", required=True, help="output path without extension")
    m.add_argument("--command", default="", help="the profiled command, for the summary")
    a = ap.parse_args()
    if a.cmd == "vmem":
        text = (f"# Vector-memory pipeline per kernel\n\nCommand per pass: `{a.command}` with "
                f"`--pmc` set to each of:\n\n" +
                "".join(f"- {k}: `{v}`\n" for k, v in VMEM_PASSES.items()) +
                "\nPer CU and cycle (GRBM_GUI_ACTIVE over the 8 XCDs, 256 CUs): vector-L1 cache "
                "accesses (tag lookups), vector load instructions, and the busy and stall "
                "shares of the texture-address (TA) and texture-data (TD) units and the L1.\n\n"
                + vmem_table(a.base))
        with open(a.out + ".md", "w") as fh:
            fh.write(text)
        print(text)
        return
    if a.cmd == "counters":
        text = (f"# Counter summary per kernel\n\nCommand per pass: `{a.command}` with "
                f"`--pmc` set to each of:\n\n" +
                "".join(f"- {k}: `{v}`\n" for k, v in COUNTER_PASSES.items()) +
                "\nDuration from GRBM_GUI_ACTIVE at 2.4 GHz; VALU busy = VALU instructions x 4 "
                "cycles over 256 CUs x 4 SIMDs; LDS bank-conflict cycles as a share of the "
                "kernel's cycles per CU.\n\n" + counter_table(a.base))
        with open(a.out + ".md", "w") as fh:
            fh.write(text)
        clean = {k: {n: (None if v != v else v) for n, v in c.items()}       # NaN -> null
                 for k, c in counter_values(a.base).items()}
        with open(a.out + ".json", "w") as fh:
            json.dump({"command": a.command, "kernels": clean}, fh, indent=1)
        print(text)
        return
    if a.cmd == "traffic":
        s = summarize(a.fetch_dir, a.write_dir)
        with open(a.out + ".json", "w") as fh:
            json.dump(s, fh, indent=1)
        text = to_markdown(s, a.command)
    else:
        text = f"# {a.title}\n\nCommand: `{a.command}`\n\n" + kernel_stats(a.run_dir)
    with open(a.out + ".md", "w") as fh:
        fh.write(text)
    print(text)


if __name__ == "__main__":
    main()
