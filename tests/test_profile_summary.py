"""Host logic of dmrgpy_amd.profile_summary (CPU): kernel roles and
categories from rocprofv3 kernel names, grouping of dispatches into H_eff
applies, counter calibration and the kernel-stats tables."""

import csv
import json
import subprocess
import sys

from dmrgpy_amd import profile_summary as ps


def test_roles_and_categories():
    assert ps._role("void dmrg::convert_k<float>(double const*, float*, long long)") == "to_fp32"
    assert ps._role("void dmrg::split_transpose_k<double>(double const*, int)") == "to_fp32"
    assert ps._role("void dmrg::convert_back_k<float>(float const*, double*, long long)") == "to_fp64"
    assert ps._role("void dmrg::gemm_fast<float, 4, 4, false, false, 1>(float const*)") == "L-GEMM"
    assert ps._role("void dmrg::gemm_fast<float, 4, 4, false, false, 2, double>(float const*)") == "R-GEMM"
    assert ps._role("void dmrg::gemm_b3nt_k<4, 4, 2>(__bf16 const*)") == "R-GEMM"
    assert ps._role("_ZN4dmrg16smallop_sparse_kIfLb0EEEvPKT_") == "small-op"
    assert ps._role("void dmrg::gemm_fast<double, 4, 2, false, true, 4>(double const*)") == "gemm-other"
    assert ps._category("dmrg::trd_symv_k(double*, int)") == "trd"
    assert ps._category("dmrg::trc_col_k(double const*, int)") == "trd"
    assert ps._category("void rocsolver::v33001::stedc_solve_kernel<double>(int)") == "eigensolver"
    assert ps._category("Cijk_Alik_Bljk_DB_MT128x64x16") == "eigensolver"
    assert ps._category("void dmrg::gemm_fast<float, 2, 2, true, false, 3>(float const*)") == "env"
    assert ps._category("void dmrg::dot_block_k<false>(double const*)") == "krylov"
    assert ps._category("__amd_rocclr_copyBuffer") == "copy"


def _write_trace(path, names, counter=None, values=None):
    with open(path, "w", newline="") as f:
        fields = ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"]
        if counter:
            fields += ["Counter_Name", "Counter_Value"]
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        t = 0
        for i, nm in enumerate(names):
            row = {"Dispatch_Id": i, "Kernel_Name": nm, "Start_Timestamp": t, "End_Timestamp": t + 1000 * (i + 1)}
            if counter:
                row["Counter_Name"] = counter
                row["Counter_Value"] = values[i]
            w.writerow(row)
            t += 100000


APPLY = ["void dmrg::convert_k<float>(x)", "void dmrg::gemm_fast<float, 4, 4, false, false, 1>(x)",
         "void dmrg::smallop_sparse_k<float, false>(x)", "void dmrg::gemm_fast<float, 4, 4, false, false, 2, double>(x)"]


def test_group_applies(tmp_path):
    names = ["void dmrg::gemm_fast<double, 2, 2, false, true, 4>(x)"] + APPLY + APPLY + ["dmrg::trd_col_k(x)"]
    p = tmp_path / "t.csv"
    _write_trace(p, names)
    ap = ps.group_applies(ps.read_dispatches(str(p)))
    assert len(ap) == 2
    assert [d[1] for d in ap[0]] == ["to_fp32", "L-GEMM", "small-op", "R-GEMM"]
    dur = ps.per_role(ap, 3)
    assert set(dur) == {"to_fp32", "L-GEMM", "small-op", "R-GEMM"}


def test_calibration_and_cli(tmp_path):
    cal = tmp_path / "cal.csv"
    _write_trace(cal, ["__amd_rocclr_copyBuffer"] * 3, "FETCH_SIZE", [512.0, 512.0, 512.0])
    assert ps.calibration(str(cal), "FETCH_SIZE", 1 << 20) == 2048.0
    stats = tmp_path / "k.csv"
    with open(stats, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs",
                                          "MaxNs", "StdDev"])
        w.writeheader()
        w.writerow({"Name": "dmrg::trd_symv_k(double*)", "Calls": 10, "TotalDurationNs": 3000, "AverageNs": 300,
                    "Percentage": 75, "MinNs": 1, "MaxNs": 1, "StdDev": 0})
        w.writerow({"Name": "void dmrg::gemm_fast<float, 4, 4, false, false, 1>(x)", "Calls": 1,
                    "TotalDurationNs": 1000, "AverageNs": 1000, "Percentage": 25, "MinNs": 1, "MaxNs": 1,
                    "StdDev": 0})
    out = subprocess.check_output([sys.executable, "-m", "dmrgpy_amd.profile_summary", "--stats", str(stats)],
                                  text=True)
    assert "| 75.0% | 0.000 | own Householder tridiagonalisation" in out
    assert "H_eff GEMMs" in out
    trace = tmp_path / "f.csv"
    _write_trace(trace, APPLY * 2, "FETCH_SIZE", [1.0, 2.0, 3.0, 4.0] * 2)
    out = subprocess.check_output([sys.executable, "-m", "dmrgpy_amd.profile_summary", "--applies", "2",
                                   "--fetch", str(trace), "--write", str(trace)], text=True)
    res = json.loads(out)
    assert abs(res["fetch_bytes"]["L-GEMM"] - 2.0 * 1024) < 1e-9
