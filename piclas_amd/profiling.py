"""Summaries of rocprofv3 CSV output for the particle stage.

  python -m piclas_amd.profiling stats  <kernel_stats.csv>
      per-kernel calls / average / share table (from --kernel-trace --stats)
  python -m piclas_amd.profiling traffic <fetch.csv> <write.csv> <copy_fetch.csv> <copy_write.csv> [copy_bytes]
      HBM bytes per dispatch of every kernel from FETCH_SIZE / WRITE_SIZE
      (KiB per dispatch), each scaled by the factor that makes the same
      counter read the known size of a plain device copy (hbm_copy_probe).
  python -m piclas_amd.profiling issue <counters.csv> [kernel-substring ...]
      per dispatch: instruction counts (SQ_INSTS_VALU, _MFMA, _LDS, _SALU,
      _VMEM_RD), SQ_VALU_MFMA_BUSY_CYCLES and the share of the SIMD cycles
      that VALU issue (4 cycles per wave instruction on a 16-lane SIMD) and
      the matrix cores take together; GRBM_GUI_ACTIVE counts the cycles of
      each of the 8 XCDs (MI355X: 8 XCDs, 1024 SIMDs).
  python -m piclas_amd.profiling perwave <counters.csv> <kernel> [kernel ...]
      the wave-state counters of the named kernels (exact short names) per
      wave: each counter summed over the kernel's dispatches, divided by
      the waves they launched (Grid_Size / 64 per dispatch).

All print Markdown tables.
"""
import csv
import sys
from collections import OrderedDict


def short(name):
    """Kernel name without namespace and argument list."""
    n = name.strip('"')
    n = n.replace("void ", "").replace("(anonymous namespace)::", "")
    depth, out = 0, []
    for ch in n:
        if ch == "(" and depth == 0:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out)


def kernel_stats(path, top=12):
    rows = list(csv.DictReader(open(path)))
    lines = ["| kernel | calls | avg (us) | min (us) | max (us) | share |", "|---|---|---|---|---|---|"]
    for r in rows[:top]:
        lines.append("| `%s` | %s | %.1f | %.1f | %.1f | %.2f %% |" % (
            short(r["Name"]), r["Calls"], float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3,
            float(r["MaxNs"]) / 1e3, float(r["Percentage"])))
    return "\n".join(lines)


def counter_per_kernel(path):
    out = OrderedDict()
    for r in csv.DictReader(open(path)):
        out.setdefault(short(r["Kernel_Name"]), []).append(float(r["Counter_Value"]))
    return out


def calibration(copy_csv, copy_bytes):
    """bytes moved per counter unit for the plain copy (largest dispatches)."""
    vals = [v for k, vs in counter_per_kernel(copy_csv).items() if "copyBuffer" in k for v in vs]
    kib = max(vals)
    return copy_bytes / (kib * 1024.0), kib


def traffic(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, copy_bytes=1 << 30):
    ff, fk = calibration(copy_fetch_csv, copy_bytes)
    fw, wk = calibration(copy_write_csv, copy_bytes)
    F = counter_per_kernel(fetch_csv)
    W = counter_per_kernel(write_csv)
    lines = ["Calibration: a %d-byte device copy reads FETCH_SIZE = %.0f KiB (factor %.3f) and WRITE_SIZE = %.0f KiB "
             "(factor %.3f).\n" % (copy_bytes, fk, ff, wk, fw),
             "| kernel | dispatches | FETCH_SIZE (KiB) | WRITE_SIZE (KiB) | HBM read (MB) | HBM write (MB) | traffic (MB) |",
             "|---|---|---|---|---|---|---|"]
    for k in F:
        f = sum(F[k]) / len(F[k])
        w = sum(W.get(k, [0.0])) / max(len(W.get(k, [0.0])), 1)
        rb, wb = f * 1024 * ff, w * 1024 * fw
        lines.append("| `%s` | %d | %.0f | %.0f | %.1f | %.1f | %.1f |" % (k, len(F[k]), f, w, rb / 1e6, wb / 1e6,
                                                                        (rb + wb) / 1e6))
    return "\n".join(lines)


def issue(path, names=(), xcds=8, simds=1024):
    agg, disp = OrderedDict(), OrderedDict()
    for r in csv.DictReader(open(path)):
        k = short(r["Kernel_Name"])
        if names and not any(n in k for n in names):
            continue
        agg.setdefault(k, {}).setdefault(r["Counter_Name"], 0.0)
        agg[k][r["Counter_Name"]] += float(r["Counter_Value"])
        disp.setdefault(k, set()).add(r["Dispatch_Id"])
    lines = ["| kernel | dispatches | cycles per XCD | VALU | MFMA | LDS | SALU | VMEM rd | MFMA busy cycles | "
             "VALU+MFMA share of SIMD cycles |", "|---|---|---|---|---|---|---|---|---|---|"]
    for k, c in agg.items():
        n = len(disp[k])
        g = lambda name: c.get(name, 0.0) / n
        cyc = g("GRBM_GUI_ACTIVE") / xcds
        share = (4.0 * g("SQ_INSTS_VALU") + g("SQ_VALU_MFMA_BUSY_CYCLES")) / (cyc * simds) if cyc > 0 else 0.0
        lines.append("| `%s` | %d | %.4g | %.4g | %.4g | %.4g | %.4g | %.4g | %.4g | %.2f |" % (
            k, n, cyc, g("SQ_INSTS_VALU"), g("SQ_INSTS_MFMA"), g("SQ_INSTS_LDS"), g("SQ_INSTS_SALU"),
            g("SQ_INSTS_VMEM_RD"), g("SQ_VALU_MFMA_BUSY_CYCLES"), share))
    return "\n".join(lines)


PERWAVE = ("SQ_WAVE_CYCLES", "SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_LDS", "SQ_ACTIVE_INST_MISC",
           "SQ_WAIT_INST_ANY", "SQ_WAIT_INST_LDS", "SQ_LDS_BANK_CONFLICT")


def perwave(path, names, wave=64):
    agg, waves = OrderedDict((k, {}) for k in names), {k: {} for k in names}
    for r in csv.DictReader(open(path)):
        k = short(r["Kernel_Name"])
        if k not in agg:
            continue
        agg[k][r["Counter_Name"]] = agg[k].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        waves[k][r["Dispatch_Id"]] = int(r["Grid_Size"]) // wave

    def fmt(v):
        return "%.3e" % v if v >= 1e4 else ("%d" % round(v) if v >= 10 else "%.1f" % v)

    lines = ["| kernel | " + " | ".join(PERWAVE) + " |", "|" + "---|" * (len(PERWAVE) + 1)]
    for k, c in agg.items():
        w = sum(waves[k].values())
        if w:
            lines.append("| `%s` | " % k + " | ".join(fmt(c.get(n, 0.0) / w) for n in PERWAVE) + " |")
    return "\n".join(lines)


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        print(kernel_stats(argv[1]))
    elif len(argv) >= 2 and argv[0] == "issue":
        print(issue(argv[1], tuple(argv[2:])))
    elif len(argv) >= 3 and argv[0] == "perwave":
        print(perwave(argv[1], tuple(argv[2:])))
    elif len(argv) >= 5 and argv[0] == "traffic":
        nb = int(argv[5]) if len(argv) > 5 else (1 << 30)
        print(traffic(argv[1], argv[2], argv[3], argv[4], nb))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
