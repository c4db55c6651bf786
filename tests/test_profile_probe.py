"""The rocprofv3 summarizers behind the committed profiles, on synthetic CSV input.

Counter values are chosen so each derived column has a hand-computed answer: a kernel
lasting 1000 cycles (GRBM_GUI_ACTIVE summed over 8 XCDs), issuing 256 000 wave64 VALU
instructions (250 per SIMD of 1024, 4 cycles each: 100 % issue) with 128 000 VALU-active
CU cycles (500 per CU: 50 % busy); and a copy calibration where FETCH_SIZE reads half
the copied bytes in KiB and WRITE_SIZE reads them exactly.
"""

import csv

from caustica_amd import profile_probe as pp


def _write(path, rows):
    cols = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=cols)
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(cols, r)))


def test_summarize_sq_derived_columns(tmp_path):
    k = "void cst::k_zu<128, 2>(cst::StepArgs)"
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    gui = 8 * 1000.0  # 1000 cycles on each of 8 XCDs
    _write(a / "run_counter_collection.csv", [
        (k, "GRBM_GUI_ACTIVE", gui), (k, "SQ_WAVES", 4096), (k, "SQ_WAVE_CYCLES", 256 * 1000 * 16 / 4),
        (k, "SQ_ACTIVE_INST_VALU", 256 * 500), (k, "SQ_INSTS_LDS", 4096 * 10),
        (k, "SQ_LDS_BANK_CONFLICT", 50), (k, "SQ_ACTIVE_INST_LDS", 500),
    ])
    _write(b / "run_counter_collection.csv", [(k, "GRBM_GUI_ACTIVE", gui), (k, "SQ_INSTS_VALU", 256 * 4 * 250)])
    out = pp.summarize_sq(str(a), str(b))
    row = [ln for ln in out.splitlines() if "k_zu<128, 2>" in ln][0]
    cells = [c.strip() for c in row.strip("|").split("|")]
    # kernel, waves, VALU/wave, VALU issue %, VALU busy %, waves/CU, LDS/wave, conflict %
    assert cells[1] == "4096"
    assert cells[2] == "62"     # 256 000 / 4096 = 62.5, printed to the unit (half to even)
    assert cells[3] == "100"    # 250 instructions x 4 cycles per SIMD over 1000 cycles
    assert cells[4] == "50"     # 500 active cycles per CU over 1000
    assert cells[5] == "16.0"   # 4 x 1 024 000 quad-cycle wave counts / 256 CUs / 1000
    assert cells[6] == "10"     # 40 960 LDS instructions / 4096 waves
    assert cells[7] == "10"     # 50 conflict cycles / 500 LDS-active cycles


def test_summarize_traffic_scales_by_calibration(tmp_path):
    f, w = tmp_path / "f", tmp_path / "w"
    f.mkdir()
    w.mkdir()
    cal = "void cst::k_diag_copy(float4 const*, float4*, long)"
    k = "void cst::k_xgrad<256>(cst::StepArgs)"
    # calibration: 256 MiB read + 256 MiB written; FETCH_SIZE counts half the bytes in KiB
    _write(f / "run_counter_collection.csv", [(cal, "FETCH_SIZE", pp.CAL_BYTES / 2048), (k, "FETCH_SIZE", 1000.0)])
    _write(w / "run_counter_collection.csv", [(cal, "WRITE_SIZE", pp.CAL_BYTES / 1024), (k, "WRITE_SIZE", 500.0)])
    out = pp.summarize_traffic(str(f), str(w), "wl")
    line = [ln for ln in out.splitlines() if ln.startswith("traffic_bytes_per_launch xgrad")][0]
    assert int(line.split()[-1]) == 1000 * 2048 + 500 * 1024
