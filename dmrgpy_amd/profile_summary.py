"""Summarise rocprofv3 CSV output for the two-site H_eff apply.

    python -m dmrgpy_amd.profile_summary --applies K [--trace T.csv]
        [--fetch F.csv --write W.csv --cal-fetch CF.csv --cal-write CW.csv
         --cal-bytes B]
    python -m dmrgpy_amd.profile_summary --stats KERNEL_STATS.csv [--top 20]

An H_eff apply in reduced precision is five consecutive dispatches on the
engine's stream: FP64->FP32 conversion of theta, the L-GEMM (kernel tag 1),
the small operator, the R-GEMM (tag 2) and the conversion of y back to
FP64 (FP64 applies have only the three middle ones).  The rocprofv3 CSVs
(``--kernel-trace`` output, or ``--pmc`` counter collection, filtered with
``--kernel-include-regex 'gemm_fast|smallop|convert'``) are grouped into
applies in dispatch order and the last K applies -- the timed sweep of a
``bench.py --no-fp64 --no-cpu`` run, whose H_eff count the bench prints --
are summarised: average duration per role (kernel trace) and HBM bytes per
apply (FETCH_SIZE / WRITE_SIZE in KiB, scaled by the bytes a device copy of
known size reports for the same counters).  Prints Markdown.
"""

import argparse
import csv
import json

ROLES = ("to_fp32", "L-GEMM", "small-op", "R-GEMM", "to_fp64")


def _role(name):
    if name.startswith("void dmrg::convert_k") or name.startswith("void dmrg::split_transpose_k<double>"):
        return "to_fp32"
    if name.startswith("void dmrg::convert_back_k"):
        return "to_fp64"
    if "smallop" in name:   # the split-output variant's name reaches the CSV mangled
        return "small-op"
    if name.startswith("void dmrg::gemm"):
        tag = _gemm_tag(name)
        return {"1": "L-GEMM", "2": "R-GEMM"}.get(tag, "gemm-other")
    return "other"


def _gemm_tag(name):
    """The tag template argument of a dmrg GEMM kernel name: the last integer
    argument (an output-type argument may follow it)."""
    args = [a.strip() for a in name.split("<", 1)[1].split(">(")[0].split(",")]
    ints = [a for a in args if a.lstrip("-").isdigit()]
    return ints[-1] if ints else ""


def read_dispatches(path, counter=None):
    """[(dispatch_id, role, name, duration_ns, value)] in dispatch order."""
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if counter is not None and r.get("Counter_Name") != counter:
                continue
            dur = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            val = float(r["Counter_Value"]) if "Counter_Value" in r else 0.0
            out.append((int(r["Dispatch_Id"]), _role(r["Kernel_Name"]), r["Kernel_Name"], dur, val))
    out.sort(key=lambda t: t[0])
    return out


def group_applies(disp):
    """Applies as lists of dispatches: an L-GEMM, the small-op and R-GEMM
    right after it, plus an adjacent to_fp32 before / to_fp64 after."""
    applies = []
    i = 0
    while i < len(disp):
        if disp[i][1] == "L-GEMM" and i + 2 < len(disp) and disp[i + 1][1] == "small-op" \
                and disp[i + 2][1] == "R-GEMM":
            grp = [disp[i], disp[i + 1], disp[i + 2]]
            if i > 0 and disp[i - 1][1] == "to_fp32":
                grp.insert(0, disp[i - 1])
            if i + 3 < len(disp) and disp[i + 3][1] == "to_fp64":
                grp.append(disp[i + 3])
            applies.append(grp)
            i += 3
        else:
            i += 1
    return applies


def per_role(applies, field):
    tot = {}
    for grp in applies:
        for d in grp:
            tot[d[1]] = tot.get(d[1], 0.0) + d[field]
    return {k: v / len(applies) for k, v in tot.items()}


def calibration(path, counter, nbytes):
    """bytes per counter KiB from the copy dispatches of a calibration run."""
    vals = [d[4] for d in read_dispatches(path, counter) if "copyBuffer" in d[2] or "copy" in d[2].lower()]
    if not vals:
        raise SystemExit("no copy dispatch in " + path)
    return nbytes / (sum(vals) / len(vals))


CATEGORIES = (
    ("trd", "own Householder tridiagonalisation (real: trd_col_k, trd_symv_k, trd_fin_k; complex Hermitian: "
            "trc_col_k, trc_symv_k, trc_fin_k; its panel GEMMs count as SVD GEMMs)"),
    ("eigensolver", "rocSOLVER/rocBLAS: tridiagonal divide and conquer (stedc), back-transformation (ormtr), "
                    "syevd/heevd where used, and the library GEMMs they call"),
    ("heff", "H_eff GEMMs (L- and R-step)"),
    ("smallop", "small operator (MPO W) passes"),
    ("svd", "SVD GEMMs (Gram, factors, Loewdin, guided subspace)"),
    ("env", "environment-update GEMMs"),
    ("grad", "gradient GEMMs"),
    ("krylov", "Krylov vector kernels"),
    ("convert", "precision conversions and bf16 splits"),
    ("copy", "runtime copies and fills"),
    ("other", "other"),
)


def _category(name):
    if "trd_" in name or "trc_" in name:
        return "trd"
    if "rocsolver" in name or "rocblas" in name or name.startswith("Cijk_"):
        return "eigensolver"
    if name.startswith("void dmrg::gemm"):
        tag = _gemm_tag(name)
        return {"1": "heff", "2": "heff", "3": "env", "4": "svd", "5": "grad"}.get(tag, "other")
    if "smallop" in name:
        return "smallop"
    if any(k in name for k in ("dot_block", "sub_block", "reduce_rows", "combine_k", "scale_copy", "axpy")):
        return "krylov"
    if any(k in name for k in ("convert", "split_rows", "split_transpose")):
        return "convert"
    if name.startswith("__amd_rocclr"):
        return "copy"
    return "other"


def kernel_stats(path, top):
    """Markdown tables from a rocprofv3 --stats kernel_stats.csv."""
    rows = list(csv.DictReader(open(path)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    cat = {}
    for r in rows:
        k = _category(r["Name"])
        cat[k] = cat.get(k, 0.0) + float(r["TotalDurationNs"])
    out = ["Total kernel time {:.2f} s, {} distinct kernels.".format(tot / 1e9, len(rows)), "",
           "| share | kernel time (s) | category |", "|---:|---:|---|"]
    for k, desc in sorted(CATEGORIES, key=lambda kd: -cat.get(kd[0], 0.0)):
        if k in cat:
            out.append("| {:.1f}% | {:.3f} | {} |".format(100 * cat[k] / tot, cat[k] / 1e9, desc))
    out += ["", "Top {} kernels:".format(top), "", "| share | total (ms) | calls | avg (us) | kernel |",
            "|---:|---:|---:|---:|---|"]
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    for r in rows[:top]:
        t = float(r["TotalDurationNs"])
        out.append("| {:.2f}% | {:.1f} | {} | {:.1f} | `{}` |".format(
            100 * t / tot, t / 1e6, r["Calls"], float(r["AverageNs"]) / 1e3, r["Name"][:110]))
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", help="kernel_stats.csv of a --kernel-trace --stats run: print the category and "
                                    "top-kernel tables and exit")
    ap.add_argument("--top", type=int, default=20)
    ap.add_argument("--applies", type=int, default=0)
    ap.add_argument("--trace")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--cal-fetch")
    ap.add_argument("--cal-write")
    ap.add_argument("--cal-bytes", type=float, default=float(1 << 30))
    ap.add_argument("--mfma", help="counter CSV of a --pmc MfmaUtil run: duration-weighted MFMA utilisation "
                                   "(percent of the SIMDs' cycles busy in MFMA) per role")
    a = ap.parse_args()
    if a.stats:
        print(kernel_stats(a.stats, a.top))
        return
    if a.applies <= 0:
        raise SystemExit("--applies K is required without --stats")
    res = {}
    if a.trace:
        ap_ = group_applies(read_dispatches(a.trace))[-a.applies:]
        dur = per_role(ap_, 3)
        res["applies_found"] = len(ap_)
        res["avg_us"] = {k: dur[k] / 1e3 for k in ROLES if k in dur}
        res["avg_apply_us"] = sum(dur.values()) / 1e3
    if a.fetch and a.write:
        kf = calibration(a.cal_fetch, "FETCH_SIZE", a.cal_bytes) if a.cal_fetch else 1024.0
        kw = calibration(a.cal_write, "WRITE_SIZE", a.cal_bytes) if a.cal_write else 1024.0
        fa = group_applies(read_dispatches(a.fetch, "FETCH_SIZE"))[-a.applies:]
        wa = group_applies(read_dispatches(a.write, "WRITE_SIZE"))[-a.applies:]
        f = per_role(fa, 4)
        w = per_role(wa, 4)
        res["bytes_per_counter_kib"] = {"FETCH_SIZE": kf, "WRITE_SIZE": kw}
        res["fetch_bytes"] = {k: f[k] * kf for k in ROLES if k in f}
        res["write_bytes"] = {k: w[k] * kw for k in ROLES if k in w}
        res["traffic_bytes_per_apply"] = sum(res["fetch_bytes"].values()) + sum(res["write_bytes"].values())
    if a.mfma:
        ma = group_applies(read_dispatches(a.mfma, "MfmaUtil"))[-a.applies:]
        num, den = {}, {}
        for grp in ma:
            for d in grp:
                num[d[1]] = num.get(d[1], 0.0) + d[4] * d[3]
                den[d[1]] = den.get(d[1], 0.0) + d[3]
        res["mfma_util_percent"] = {k: num[k] / den[k] for k in ROLES if k in num and den[k] > 0}
        tot_n, tot_d = sum(num.values()), sum(den.values())
        res["mfma_util_percent_apply"] = tot_n / tot_d if tot_d > 0 else None
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
