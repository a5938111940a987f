"""The measurement helpers (CPU): the HBM traffic table built from two
rocprofv3 counter runs scaled by a plain copy's reading, the matrix-core
table from a counter run and a kernel trace, and kernel-name parsing, on
small CSV files written here in rocprofv3's column layout."""

import csv

import pytest

from plainmp_amd import roofline

COUNTER_COLS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size",
                "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
                "Accum_VGPR_Count", "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]
TRACE_COLS = ["Kind", "Agent_Id", "Queue_Id", "Stream_Id", "Thread_Id", "Dispatch_Id", "Kernel_Id", "Kernel_Name",
              "Correlation_Id", "Start_Timestamp", "End_Timestamp", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
              "Accum_VGPR_Count", "SGPR_Count", "Workgroup_Size_X", "Workgroup_Size_Y", "Workgroup_Size_Z",
              "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"]


def _write(path, cols, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=cols)
        w.writeheader()
        for r in rows:
            w.writerow({c: r.get(c, 0) for c in cols})


def _counter(name, grid, counter, value):
    return {"Kernel_Name": name, "Grid_Size": grid, "Counter_Name": counter, "Counter_Value": value}


def test_kernel_base_names():
    assert roofline._kernel_base("void pmp::edges_kernel<false>(pmp::KinArgs<float>, float)") == "edges_kernel<false>"
    assert roofline._kernel_base("pmp::(anonymous namespace)::tiles_kernel(double const*, long)") == "tiles_kernel"
    assert roofline._kernel_base("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"
    assert roofline._kernel_base("void pmp::pair_kernel<8, false>(pmp::KinArgs<float>)") == "pair_kernel<8, false>"


def test_traffic_table_scales_by_the_copy(tmp_path):
    # the copy of 1,024 MiB reads FETCH_SIZE 524,288 KiB (half) and WRITE_SIZE 1,048,576 KiB
    copy = "__amd_rocclr_copyBuffer"
    k = "void pmp::pair_kernel<8, false>(pmp::KinArgs<float>)"
    fetch = [_counter(copy, 131072, "FETCH_SIZE", 524288.0), _counter(k, 65536, "FETCH_SIZE", 100.0),
             _counter(k, 65536, "FETCH_SIZE", 300.0), _counter(k, 196608, "FETCH_SIZE", 400.0)]
    write = [_counter(copy, 131072, "WRITE_SIZE", 1048576.0), _counter(k, 65536, "WRITE_SIZE", 1000.0),
             _counter(k, 65536, "WRITE_SIZE", 1000.0), _counter(k, 196608, "WRITE_SIZE", 2000.0)]
    _write(tmp_path / "f.csv", COUNTER_COLS, fetch)
    _write(tmp_path / "w.csv", COUNTER_COLS, write)
    t = roofline.traffic_table(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), 1024)
    assert t["_calibration"]["fetch_factor"] == pytest.approx(2.0)
    assert t["_calibration"]["write_factor"] == pytest.approx(1.0)
    rec = t["pair_kernel<8, false>"]
    assert rec["grids"]["65536"]["launches"] == 2
    assert rec["grids"]["65536"]["bytes_per_launch"] == (200 * 2 + 1000) * 1024  # mean fetch 200 KiB x 2
    assert rec["bytes_per_launch"] == (400 * 2 + 2000) * 1024  # the largest grid
    assert t["_calibration"]["sources_digest"] == roofline.sources_digest()


def test_traffic_table_needs_the_copy(tmp_path):
    k = "void pmp::pair_kernel<8, false>(pmp::KinArgs<float>)"
    _write(tmp_path / "f.csv", COUNTER_COLS, [_counter(k, 65536, "FETCH_SIZE", 100.0)])
    _write(tmp_path / "w.csv", COUNTER_COLS, [_counter(k, 65536, "WRITE_SIZE", 100.0)])
    with pytest.raises(ValueError):
        roofline.traffic_table(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), 1024)


def test_matrix_core_table(tmp_path):
    k = "void pmp::cloud_min_kernel<true>(pmp::CloudTile const*, float const*)"
    other = "void pmp::edges_kernel<false>(pmp::KinArgs<float>)"
    rows = [_counter(k, 6291456, "SQ_INSTS_VALU_MFMA_F32", 60000.0), _counter(k, 6291456, "SQ_INSTS_VALU_MFMA_F32", 64000.0),
            _counter(k, 6291456, "SQ_VALU_MFMA_BUSY_CYCLES", 2000000.0),
            _counter(other, 262144, "SQ_INSTS_VALU_MFMA_F32", 0.0)]
    _write(tmp_path / "m.csv", COUNTER_COLS, rows)
    trace = [{"Kernel_Name": k, "Start_Timestamp": 1000, "End_Timestamp": 101000, "Grid_Size_X": 6291456,
              "Grid_Size_Y": 1, "Grid_Size_Z": 1},
             {"Kernel_Name": k, "Start_Timestamp": 200000, "End_Timestamp": 310000, "Grid_Size_X": 6291456,
              "Grid_Size_Y": 1, "Grid_Size_Z": 1},
             {"Kernel_Name": other, "Start_Timestamp": 0, "End_Timestamp": 5000, "Grid_Size_X": 262144,
              "Grid_Size_Y": 1, "Grid_Size_Z": 1}]
    _write(tmp_path / "t.csv", TRACE_COLS, trace)
    m = roofline.matrix_core_table(str(tmp_path / "m.csv"), str(tmp_path / "t.csv"))
    assert list(m) == ["cloud_min_kernel<true>"]  # kernels without MFMA are left out
    rec = m["cloud_min_kernel<true>"]["6291456"]
    assert rec["mfma_f32_insts"] == pytest.approx(62000.0)
    assert rec["mfma_busy_cycles"] == pytest.approx(2000000.0)
    assert rec["duration_us"] == pytest.approx(105.0)  # (100 + 110) / 2 µs
    assert rec["launches"] == 2
