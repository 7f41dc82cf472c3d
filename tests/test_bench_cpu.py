"""Host-side figures of bench.py (no GPU): the candidate count of the cell
walk that the VALU-per-candidate figure divides by, against a brute-force
count with the oracle's HashCell3D (pinned by its known answers); the
per-wave counter summary and the calibrated HBM traffic on synthetic
rocprofv3 databases."""
import numpy as np
import pytest

import bench
from harness import ORC


def _brute_candidates(x, cell, n):
    c = np.floor(x / np.float32(cell)).astype(np.int64)
    key = lambda a, b, d: ORC.orc_hash_cell(int(a), int(b), int(d)) % n
    counts = np.bincount([key(*cc) for cc in c], minlength=n)
    total = 0
    for cx, cy, cz in c:
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    total += int(counts[key(cx + dx, cy + dy, cz + dz)])
    return total / len(x)


@pytest.mark.parametrize("n, lo, hi, cell", [
    (600, 0.0, 0.5, 0.05),     # one 50-block of cells
    (600, -2.0, 3.0, 0.05),    # cells below -25 (wrapping) and across 50-blocks
    (37, -0.2, 0.2, 0.04),     # tiny table: stencil cells share keys (buckets visited twice)
])
def test_candidates_per_particle_matches_brute_force(n, lo, hi, cell):
    rng = np.random.default_rng(n)
    x = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    assert bench.candidates_per_particle(x, cell, n) == pytest.approx(_brute_candidates(x, cell, n), rel=0, abs=0)


def _counter_db(path, rows):
    import sqlite3

    con = sqlite3.connect(str(path))
    con.execute("create table counters_collection (kernel_name text, counter_name text, value real, grid_size int)")
    con.executemany("insert into counters_collection values (?, ?, ?, ?)", rows)
    con.commit()
    con.close()
    return str(path)


def test_counter_summary_divides_by_waves_and_lane_cycles(tmp_path):
    # two launches of the density kernel: per-launch means, per-wave figures, and the
    # active-lane fraction SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU)
    name = "void vfs::(anonymous namespace)::k_density<128, true, false>(vfs::StepBuffers, ...)"
    rows = [(name, "SQ_WAVES", 100.0, 0), (name, "SQ_WAVES", 100.0, 0),
            (name, "SQ_INSTS_VALU", 8000.0, 0), (name, "SQ_INSTS_VALU", 12000.0, 0),
            (name, "SQ_ACTIVE_INST_VALU", 1000.0, 0), (name, "SQ_ACTIVE_INST_VALU", 1000.0, 0),
            (name, "SQ_THREAD_CYCLES_VALU", 48000.0, 0), (name, "SQ_THREAD_CYCLES_VALU", 48000.0, 0),
            ("k_other", "SQ_WAVES", 5.0, 0)]
    text = bench.summarize_counters([_counter_db(tmp_path / "a.db", rows)])
    assert "k_density: 100 waves per launch (mean of 2 launches)" in text
    line = [ln for ln in text.splitlines() if "SQ_INSTS_VALU" in ln][0]
    assert line.split()[-1] == "100.0"  # (8000 + 12000) / 2 / 100 waves
    assert "lanes active per VALU instruction: 0.750" in text  # 48000 / (64 * 1000)
    assert "k_other" not in text


def test_traffic_summary_scales_by_the_calibration_copy(tmp_path):
    # the 256 MiB copy reads FETCH_SIZE 131072 KiB (half the bytes) and WRITE_SIZE 262144 KiB
    mib = bench.CALIBRATION_MIB
    kib = mib * 1024.0
    copy = "__amd_rocclr_copyBuffer"
    dens = "void vfs::(anonymous namespace)::k_density<128, true, false>(...)"
    fetch = _counter_db(tmp_path / "f.db", [(copy, "FETCH_SIZE", kib / 2, 1), (dens, "FETCH_SIZE", 1000.0, 2),
                                            (dens, "FETCH_SIZE", 3000.0, 2)])
    write = _counter_db(tmp_path / "w.db", [(copy, "WRITE_SIZE", kib, 1), (dens, "WRITE_SIZE", 500.0, 2),
                                            (dens, "WRITE_SIZE", 500.0, 2)])
    text = bench.summarize_pmc(fetch, write)
    assert "FETCH_SIZE x 2.0000" in text and "WRITE_SIZE x 1.0000" in text
    row = [ln for ln in text.splitlines() if ln.startswith("k_density")][0].split()
    assert row[1] == "2" and float(row[2]) == 2000.0 and float(row[3]) == 500.0
    # (2000 * 2 + 500 * 1) KiB per launch = 4.608 MB
    assert float(row[4]) == pytest.approx((2000.0 * 2.0 + 500.0) * 1024.0 / 1e6, abs=0.05)


def test_counter_summary_memory_path_fractions(tmp_path):
    # one launch: kernel cycles from GRBM_GUI_ACTIVE over 8 XCDs, per-CU busy fractions of
    # the 256-CU sums, VALU busy from 4-cycle units over 1024 SIMDs, the L1 hit rate
    name = "void vfs::(anonymous namespace)::k_forces_integrate<128, false>(vfs::StepBuffers, ...)"
    cyc = 1000.0
    rows = [(name, "SQ_WAVES", 64.0, 0), (name, "GRBM_GUI_ACTIVE", 8 * cyc, 0),
            (name, "SQ_ACTIVE_INST_VALU", 0.5 * cyc * 1024 / 4, 0),
            (name, "TD_TD_BUSY_sum", 0.8 * cyc * 256, 0), (name, "TD_TC_STALL_sum", 0.4 * cyc * 256, 0),
            (name, "SQ_INSTS_VMEM_RD", 0.25 * cyc * 256 / 16, 0),
            (name, "TCP_TOTAL_CACHE_ACCESSES_sum", 1000.0, 0), (name, "TCP_TCC_READ_REQ_sum", 200.0, 0)]
    text = bench.summarize_counters([_counter_db(tmp_path / "m.db", rows)])
    assert "kernel cycles (GRBM_GUI_ACTIVE / 8 XCDs): 1000" in text
    assert "VALU busy (4 x SQ_ACTIVE_INST_VALU / 1024 SIMDs / kernel cycles): 0.500" in text
    assert "TD_TD_BUSY_sum per CU / kernel cycles: 0.800" in text
    assert "TD_TC_STALL_sum per CU / kernel cycles: 0.400" in text
    assert "per CU / kernel cycles: 0.250" in text  # 16 TD cycles per VMEM read instruction
    assert "L1 hit rate (1 - TCP_TCC_READ_REQ / TCP_TOTAL_CACHE_ACCESSES): 0.800" in text


def test_stats_by_grid_separates_scene_sizes(tmp_path):
    """summarize_stats_by_grid groups a --kernel-trace database's dispatches
    by kernel and grid size (durations in ns, printed in ms)."""
    import sqlite3

    db = tmp_path / "k.db"
    con = sqlite3.connect(db)
    con.execute("create table kernels (name text, grid_x integer, duration integer)")
    dens = "void vfs::(anonymous namespace)::k_density<128, true, false>(vfs::StepBuffers, ...)"
    rows = [(dens, 75776, 80000)] * 3 + [(dens, 40960, 60000), (dens, 40960, 62000)]
    con.executemany("insert into kernels values (?, ?, ?)", rows)
    con.commit()
    lines = bench.summarize_stats_by_grid(str(db)).splitlines()
    assert lines[1].split() == ["k_density<128,", "true,", "false>", "75776", "3", "0.2", "0.080"]
    assert lines[2].split() == ["k_density<128,", "true,", "false>", "40960", "2", "0.1", "0.061"]
