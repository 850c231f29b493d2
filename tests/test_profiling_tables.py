"""The rocprofv3 summaries behind profiles/: kernel names shortened, counters
summed per kernel and divided per dispatch or per wave, on a hand-made
counter file whose sums are known."""
import csv

from piclas_amd import profiling as P

FIELDS = ["Dispatch_Id", "Grid_Size", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _write(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(FIELDS, r)))


def test_short_names():
    assert P.short('void (anonymous namespace)::k_a<2, false>(pic::DevMesh, int const*)') == "k_a<2, false>"
    assert P.short("k_b(int)") == "k_b"


def test_perwave_divides_by_launched_waves(tmp_path):
    f = tmp_path / "c.csv"
    ka = "void (anonymous namespace)::k_a<2, false>(pic::DevMesh)"
    rows = []
    # two dispatches of k_a: 128 and 256 lanes -> 2 + 4 = 6 waves
    for d, grid, scale in ((1, 128, 1.0), (2, 256, 2.0)):
        for n in P.PERWAVE:
            rows.append((d, grid, ka, n, 600.0 * scale))
    rows.append((3, 64, "k_other(int)", "SQ_WAVE_CYCLES", 1e9))
    _write(f, rows)
    out = P.perwave(str(f), ("k_a<2, false>",)).splitlines()
    assert len(out) == 3 and out[2].startswith("| `k_a<2, false>` |")
    vals = [c.strip() for c in out[2].strip("|").split("|")[1:]]
    assert vals == ["300"] * len(P.PERWAVE)   # (600 + 1200) / 6


def test_issue_share(tmp_path):
    f = tmp_path / "i.csv"
    k = "k_c(int)"
    rows = [(1, 64, k, "GRBM_GUI_ACTIVE", 8 * 1000.0), (1, 64, k, "SQ_INSTS_VALU", 1024 * 100.0),
            (1, 64, k, "SQ_VALU_MFMA_BUSY_CYCLES", 1024 * 200.0)]
    _write(f, rows)
    line = P.issue(str(f)).splitlines()[2]
    # (4 * 100 + 200) cycles per SIMD over 1000 cycles per XCD
    assert line.rstrip().endswith("| 0.60 |")
