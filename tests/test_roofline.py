"""CPU checks of the roofline measurement helpers: parsing of rocprofv3 counter CSV files,
dispatch-by-dispatch merging, and the arithmetic of each roofline fraction."""
import os

import pytest

from backprojector_amd import roofline as rl

HEADER = ("Correlation_Id,Dispatch_Id,Agent_Id,Queue_Id,Process_Id,Thread_Id,Grid_Size,Kernel_Id,Kernel_Name,"
          "Workgroup_Size,LDS_Block_Size,Scratch_Size,VGPR_Count,Accum_VGPR_Count,SGPR_Count,Counter_Name,"
          "Counter_Value,Start_Timestamp,End_Timestamp\n")


def write_csv(d, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w") as f:
        f.write(HEADER)
        for disp, kern, name, val, t0, t1 in rows:
            f.write(f"{disp},{disp},1,1,1,1,256,3,\"{kern}\",256,0,0,64,0,32,{name},{val},{t0},{t1}\n")


def test_parse_sums_rows_and_keeps_launch_order(tmp_path):
    k = "void bpx::confocal_lds_kernel<bpx::FusedF32, 8, 8>(bpx::ConfocalArgs)"
    write_csv(str(tmp_path), [(1, k, "SQ_INSTS_VALU", 10, 100, 1100), (1, k, "SQ_INSTS_VALU", 5, 100, 1100),
                              (2, "void bpx::stage_rows_kernel(x)", "SQ_INSTS_VALU", 7, 0, 10),
                              (3, k, "SQ_INSTS_VALU", 30, 2000, 4000)])
    got = rl.parse_counter_csv(str(tmp_path))
    assert got[k] == [{"SQ_INSTS_VALU": 15.0, "duration_ns": 1000.0}, {"SQ_INSTS_VALU": 30.0, "duration_ns": 2000.0}]
    assert rl.signature(k) == "confocal_lds_kernel<FusedF32, 8, 8>"
    name, c = rl.find_signature(got, "confocal_lds_kernel<FusedF32")
    assert name == k and c["SQ_INSTS_VALU"] == 30.0         # the first dispatch is left out of the mean
    _, cmax = rl.find_signature(got, "confocal_lds_kernel<", pick="max")
    assert cmax["SQ_INSTS_VALU"] == 30.0
    assert rl.find_signature(got, "grid_kernel<") == (None, None)


def test_roofline_fractions():
    c = {"SQ_INSTS_VALU": 1e9, "SQ_INSTS_VALU_TRANS_F32": 1e8, "VALUBusy": 80.0, "FETCH_SIZE": 1e6,
         "WRITE_SIZE": 1e3, "SQ_LDS_BANK_CONFLICT": 1.0, "SQ_LDS_IDX_ACTIVE": 4.0}
    r = rl.kernel_roofline(c, 2.0, algorithmic_bytes=8e10)
    assert r["valu"]["frac"] == pytest.approx(1e9 * 64 / 2e-3 / (256 * 4 * 16 * 2.4e9))
    assert r["sqrt"]["frac"] == pytest.approx(4 * 1e8 * 64 / 2e-3 / (256 * 4 * 16 * 2.4e9))
    traffic = 1e6 * 1024 * rl.FETCH_FACTOR + 1e3 * 1024
    assert r["hbm"]["traffic_bytes"] == pytest.approx(traffic)
    assert r["hbm"]["frac"] == pytest.approx(traffic / 2e-3 / 8e12)
    assert r["bound"] == "valu" and r["frac"] == r["valu"]["frac"] and r["valu_busy"] == 0.8
    assert r["lds_bank_conflict_frac"] == 0.25 and r["effective_gather_Bps"] == pytest.approx(4e13)
    # a copy-like kernel: HBM binds
    h = rl.kernel_roofline({"SQ_INSTS_VALU": 1e3, "FETCH_SIZE": 4e6, "WRITE_SIZE": 4e6}, 1.0)
    assert h["bound"] == "hbm"
