"""Host tests of the rocprofv3 summariser (no GPU): the per-kernel counter
arithmetic on a synthetic counter file whose answers are known."""
import csv

from tnbp_amd import profsum

FIELDS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id",
          "Grid_Size", "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size",
          "Scratch_Size", "VGPR_Count", "Accum_VGPR_Count", "SGPR_Count", "Counter_Name",
          "Counter_Value", "Start_Timestamp", "End_Timestamp"]


def write_counters(path, dispatches):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for d, (name, grid, dur_ns, counters) in enumerate(dispatches, 1):
            for cn, cv in counters.items():
                w.writerow({"Correlation_Id": d, "Dispatch_Id": d, "Agent_Id": "Agent 1",
                            "Queue_Id": 1, "Process_Id": 1, "Thread_Id": 1, "Grid_Size": grid,
                            "Kernel_Id": 7, "Kernel_Name": name, "Workgroup_Size": 256,
                            "LDS_Block_Size": 0, "Scratch_Size": 0, "VGPR_Count": 8,
                            "Accum_VGPR_Count": 0, "SGPR_Count": 16, "Counter_Name": cn,
                            "Counter_Value": cv, "Start_Timestamp": 1000,
                            "End_Timestamp": 1000 + dur_ns})


def test_util_clock_busy_flops(tmp_path):
    # 200 us at 2.25 GHz on 8 XCDs: GUI_ACTIVE sums 8 x 450000 cycles; MFMA busy
    # 60 % of 1024 SIMDs; 17.18e9 flops = MOPS x 512
    name = "void hbp::(anonymous namespace)::k_mode_product<float, 4, 1, 18, 3>(float2 const*, int)"
    gui = 8 * 450000.0
    c = {"GRBM_GUI_ACTIVE": gui, "SQ_VALU_MFMA_BUSY_CYCLES": 0.6 * 450000 * 1024,
         "SQ_INSTS_VALU_MFMA_MOPS_F32": 17.18e9 / 512, "SQ_LDS_BANK_CONFLICT": 25.0,
         "SQ_LDS_IDX_ACTIVE": 100.0}
    path = tmp_path / "c.csv"
    write_counters(path, [(name, 524288, 200000, c), (name, 524288, 200000, c),
                          (name, 512, 5000, c)])  # the short one is filtered out
    md = profsum.util(str(path), "t", "cmd", min_us=100.0)
    row = [l for l in md.splitlines() if "k_mode_product" in l]
    assert len(row) == 1
    cells = [x.strip() for x in row[0].strip("|").split("|")]
    assert cells[0] == "`hbp::k_mode_product<float, 4, 1, 18, 3>`"
    assert cells[1:4] == ["524288", "2", "200.0"]
    assert cells[4] == "2.250"                    # GHz
    assert cells[5] == "0.600"                    # MFMA busy
    assert cells[6] == "17.18"                    # GFLOP per dispatch
    assert cells[7] == "85.9"                     # 17.18 GFLOP / 200 us
    assert cells[8] == "0.250"                    # conflicts per active LDS cycle


def test_pmc_means(tmp_path):
    name = "void hbp::(anonymous namespace)::k_scale<HIP_vector_type<double, 2u>, double>(int)"
    path = tmp_path / "p.csv"
    write_counters(path, [(name, 256, 10, {"FETCH_SIZE": 10.0}),
                          (name, 256, 10, {"FETCH_SIZE": 30.0})])
    md = profsum.pmc(str(path), "t", "cmd")
    row = [l for l in md.splitlines() if "k_scale" in l][0]
    assert "| `hbp::k_scale<double2, double>` | FETCH_SIZE | 2 | 20.0 | 10.0 | 30.0 |" == row
