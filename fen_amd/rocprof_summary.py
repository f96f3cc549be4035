"""Turn rocprofv3 CSV output into the short text summaries kept with the code.

    python -m fen_amd.rocprof_summary stats  <run_kernel_stats.csv>  --command "..."
    python -m fen_amd.rocprof_summary traffic <fetch counter_collection.csv> <write counter_collection.csv>
                                              --copy-bytes B --command "..."

`stats` lists each kernel's calls and average duration, under the library's
profiler names where the kernel is one of the step's.

`traffic` converts FETCH_SIZE and WRITE_SIZE (KiB per dispatch) into HBM
bytes per launch.  The counters are calibrated on a plain copy kernel of
known size run in the same process (bench.py --copy-probe: the HIP runtime's
device-to-device copy kernel over B bytes, reading B and writing B): factor = counter bytes / B, and a
kernel's traffic = FETCH/factor_fetch + WRITE/factor_write.  Each output line
`name bytes fetch_bytes write_bytes` is read back by bench.py.
"""
from __future__ import annotations

import argparse
import collections
import csv
import re
import sys

# demangled kernel name -> the library's profiler name (fen_profile_name)
_PATTERNS = [
    (r"project_momentum_kernel", "project_momentum"),
    (r"wall_state_kernel", "wall_state"),
    (r"(?<!adj_)momentum_march_kernel|momentum_kernel_2d", "momentum"),
    (r"project_march_kernel|project_kernel_2d", "project"),
    (r"div_r2c_kernel", "div_r2c"),
    (r"y_kernel<(?:double|float), \d+, -1, 0>", "y_fwd"),
    (r"y_kernel<(?:double|float), \d+, 1, 0>", "y_inv"),
    (r"y_kernel<(?:double|float), \d+, -?1, 1>", "y_solve"),
    (r"z_solve_kernel|z_solve_pair_kernel", "z_solve"),
    (r"thomas_c1_kernel", "thomas_c1"),
    (r"thomas_kernel|thomas_general_kernel|thomas_column_kernel", "thomas"),
    (r"vof_reconstruct_kernel", "vof_reconstruct"),
    (r"vof_sweep_kernel", "vof_sweep"),
    (r"vof_finish_kernel", "vof_finish"),
    (r"mf_fields_kernel", "mf_fields"),
    (r"mf_momentum_kernel", "mf_momentum"),
    (r"adj_momentum_(march_)?kernel", "adj_momentum"),
    (r"adj_project_kernel", "adj_project"),
    (r"adj_rhs_kernel", "adj_rhs"),
    (r"dct2_(rows|pairs)_kernel", "dct2"),
    (r"dct3_rows_kernel", "dct3"),
    (r"rfft_rows_kernel|irfft_rows_kernel", "rfft_y"),
    (r"c2r_kernel", "c2r"),
    (r"ghost_kernel", "ghost"),
    (r"checks_final_kernel|checks_kernel", "checks"),
    (r"mean_shift_kernel|apply_shift_kernel", "shift"),
    (r"copyBuffer", "hip_copy"),
]


def short_name(kernel):
    """Profiler name of a demangled kernel name; FP32 instantiations get '.f32'."""
    for rx, name in _PATTERNS:
        if re.search(rx, kernel):
            return name + (".f32" if "<float" in kernel else "")
    return kernel.split("(")[0][:60]


def stats(path, command):
    rows = list(csv.DictReader(open(path)))
    out = [f"# rocprofv3 --kernel-trace --stats: {command}",
           "# name                calls    avg_us   total_%   kernel"]
    for r in rows:
        out.append(f"{short_name(r['Name']):<20} {int(r['Calls']):>6} {float(r['AverageNs']) / 1e3:>9.1f} "
                   f"{float(r['Percentage']):>8.2f}   {r['Name'][:90]}")
    return "\n".join(out) + "\n"


def _per_kernel(path, counter):
    acc = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == counter:
            acc[r["Kernel_Name"]].append(float(r["Counter_Value"]) * 1024.0)   # KiB -> bytes
    return acc


def traffic(fetch_path, write_path, copy_bytes, command):
    fetch = _per_kernel(fetch_path, "FETCH_SIZE")
    write = _per_kernel(write_path, "WRITE_SIZE")

    def probe(acc):
        vals = [max(v) for k, v in acc.items() if short_name(k) == "hip_copy"]
        if not vals:
            raise SystemExit("no copy-probe kernel (copyBuffer) in the counter file")
        return max(vals) / copy_bytes

    ff, fw = probe(fetch), probe(write)
    out = [f"# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE: {command}",
           f"# calibration: copy probe of {copy_bytes} B read + {copy_bytes} B written; "
           f"FETCH factor {ff:.4f}, WRITE factor {fw:.4f}",
           "# name        hbm_bytes_per_launch   fetch_bytes   write_bytes   dispatches   kernel"]
    seen = set()
    for k in sorted(fetch, key=lambda k: -sum(fetch[k]) / len(fetch[k])):
        name = short_name(k)
        if name == "hip_copy" or k not in write or name in seen:
            continue
        seen.add(name)
        fb = sum(fetch[k]) / len(fetch[k]) / ff
        wb = sum(write[k]) / len(write[k]) / fw
        out.append(f"{name:<12} {fb + wb:>14.4e}   {fb:>11.4e}   {wb:>11.4e}   {len(fetch[k]):>5}   {k[:80]}")
    return "\n".join(out) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--command", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--command", required=True)
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        sys.stdout.write(stats(a.csv, a.command))
    else:
        sys.stdout.write(traffic(a.fetch_csv, a.write_csv, a.copy_bytes, a.command))


if __name__ == "__main__":
    main()
