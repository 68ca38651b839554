"""Algorithmic HBM bytes of the kernels whose roofline the benchmark reports.

Counting convention: compulsory traffic. Each distinct array a launch must read or
write counts once per launch (re-reads served by caches are not counted), except the
Gauss-Seidel sweeps, which count the per-contact gather/scatter of body velocities on
every sweep because every colour pass touches the bodies again.

Units: N bodies, E grid entries (covered cells summed over all boxes), G grid cells,
P emitted pairs, C contacts, S sweeps.
"""

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def pairs_bytes_per_body(dim):
    """k_pairs reads per body: sweep order (4), FP32 box (2D 16, 3D 24), exact FP64
    narrow data (2D q0, q1, r = 40; 3D q1, r = 32) and writes/reads one per-owner count or
    offset (4)."""
    box = 16 if dim == 2 else 24
    nar = 40 if dim == 2 else 32
    return 4 + box + nar + 4


def pairs_launch_bytes(dim, n, ncells, nentries, npairs, write_pass):
    """One k_pairs launch: per-body records, cell starts (4 per cell + 1), entries (4 each);
    the write pass adds the pair lists (8 per pair)."""
    b = n * pairs_bytes_per_body(dim) + 4 * (ncells + 1) + 4 * nentries
    if write_pass:
        b += 8 * npairs
    return b


def _record_bytes(dim):
    """Contact record with FP32 working arrays: 2D 32 B (body ids 8, normal 8, Delassus
    diagonal 4, forcing 8, mu 4); 3D 64 B (ids 8, n and s 24, arms 8, W_n and W_t 8,
    forcing 12, mu 4)."""
    return 32 if dim == 2 else 64


def gs_sweep_bytes(dim, fp32):
    """Per contact per Gauss-Seidel sweep: the record, the impulse read and written
    (2 x 4 D) and both bodies' velocity deltas read and written (2 x 2 x 4 dof, dof = 2 in
    2D, 6 in 3D): 2D 80 B, 3D 184 B. An FP64 sweep counts twice that."""
    dof = 2 if dim == 2 else 6
    b = _record_bytes(dim) + 2 * 4 * dim + 2 * 2 * 4 * dof
    return b if fp32 else 2 * b


def gs_eval_bytes(dim, fp32):
    """Residual evaluation per contact (either phase): the record, the impulse and both
    bodies read in FP64 (record + 8 D + 2 x 8 dof): 2D 80 B, 3D 184 B."""
    dof = 2 if dim == 2 else 6
    return _record_bytes(dim) + 8 * dim + 2 * 8 * dof


def gs_launch_bytes(dim, ncontacts, sweeps, evals, fp32):
    return ncontacts * (sweeps * gs_sweep_bytes(dim, fp32) + evals * gs_eval_bytes(dim, fp32))


def broadphase_bytes(dim, n, npairs):
    """The broad phase of one step as a whole (predictor with keys, binning, cell ranges,
    neighbour sweep with the exact filter, ordered pair output), counted independently of
    how it is implemented: per body 52 + 48 + 16 + 48 + 8 = 172 B in 2D and
    72 + 48 + 12 + 68 + 8 = 208 B in 3D (FP64 state read once, FP32 boxes and radii,
    three 8-byte key/id radix passes, cell ranges, per-owner count and offset), plus 8 B
    per emitted pair."""
    return (172 if dim == 2 else 208) * n + 8 * npairs


def kernel_trace_summary(trace_csv, kernel_prefix=None, last=0):
    """Per-kernel statistics from a rocprofv3 --kernel-trace CSV (one row per dispatch,
    Kernel_Name, Start_Timestamp, End_Timestamp in ns), as markdown; with kernel_prefix
    and last > 0 also the mean duration of the last `last` dispatches whose name starts
    with kernel_prefix (the benchmark's timed steps are the run's last GPU work)."""
    import csv
    rows = []
    with open(trace_csv, newline="") as f:
        for r in csv.DictReader(f):
            name = r.get("Kernel_Name") or r.get("KernelName") or ""
            t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
            rows.append((t0, name, (t1 - t0) / 1000.0))
    rows.sort()
    stats = {}
    for _, name, us in rows:
        s = stats.setdefault(name, [0, 0.0, float("inf"), 0.0])
        s[0] += 1
        s[1] += us
        s[2] = min(s[2], us)
        s[3] = max(s[3], us)
    total = sum(s[1] for s in stats.values()) or 1.0
    out = ["| kernel | calls | avg us | min us | max us | total us | % |", "|---|---|---|---|---|---|---|"]
    for name, (n, tot, mn, mx) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        short = _short(name)
        out.append("| `%s` | %d | %.1f | %.1f | %.1f | %.1f | %.1f |" % (short, n, tot / n, mn, mx, tot, 100.0 * tot / total))
    if kernel_prefix and last > 0:
        sel = [us for _, name, us in rows
               if name.replace("void ", "").replace("scx::", "").startswith(kernel_prefix)][-last:]
        if sel:
            out.append("")
            out.append("Last %d dispatches of `%s*` (the timed steps): mean %.1f us, total %.1f us." %
                       (len(sel), kernel_prefix, sum(sel) / len(sel), sum(sel)))
    return "\n".join(out)


def _short(name):
    return name.split("(")[0].replace("void ", "").replace("scx::", "")


def pmc_values(counter_csv, counter, kernel_prefix, last=0):
    """Per-dispatch values of one counter from a rocprofv3 --pmc CSV (one row per dispatch
    and counter: Dispatch_Id, Kernel_Name, Counter_Name, Counter_Value; the byte counters
    are in KiB), for the dispatches whose kernel name starts with kernel_prefix, in
    dispatch order (summed over the rows of one dispatch); last > 0 keeps the last ones."""
    import csv
    per = {}
    with open(counter_csv, newline="") as f:
        for r in csv.DictReader(f):
            if r.get("Counter_Name") != counter or not _short(r.get("Kernel_Name", "")).startswith(kernel_prefix):
                continue
            d = int(r["Dispatch_Id"])
            per[d] = per.get(d, 0.0) + float(r["Counter_Value"])
    vals = [per[d] for d in sorted(per)]
    return vals[-last:] if last > 0 else vals


def hbm_traffic(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, copy_bytes, kernel_prefix, last=0,
                copy_kernel="__amd_rocclr_copyBuffer"):
    """HBM bytes per launch of a kernel from two rocprofv3 --pmc runs (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch), each scaled by what the same counter reads for a plain
    device copy of copy_bytes (read and written once): traffic = fetch * fetch_factor +
    write * write_factor, averaged over the kernel's (last) dispatches."""
    cf = pmc_values(copy_fetch_csv, "FETCH_SIZE", copy_kernel)
    cw = pmc_values(copy_write_csv, "WRITE_SIZE", copy_kernel)
    kf = pmc_values(fetch_csv, "FETCH_SIZE", kernel_prefix, last)
    kw = pmc_values(write_csv, "WRITE_SIZE", kernel_prefix, last)
    if not (cf and cw and kf and kw):
        raise ValueError("missing dispatches in the counter files")
    ff = copy_bytes / (1024.0 * sum(cf) / len(cf))
    fw = copy_bytes / (1024.0 * sum(cw) / len(cw))
    fetch = 1024.0 * sum(kf) / len(kf) * ff
    write = 1024.0 * sum(kw) / len(kw) * fw
    return {"bytes_per_launch": fetch + write, "fetch_bytes": fetch, "write_bytes": write, "fetch_factor": ff,
            "write_factor": fw, "dispatches": [len(kf), len(kw)], "kernel": kernel_prefix,
            "method": "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs, scaled by a %d-byte device copy" % copy_bytes}


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="Kernel statistics of a rocprofv3 --kernel-trace CSV, or the HBM "
                                             "traffic of a kernel from rocprofv3 --pmc CSVs")
    ap.add_argument("trace_csv", nargs="?")
    ap.add_argument("--kernel", default=None, help="name prefix of the kernel whose last dispatches are averaged")
    ap.add_argument("--last", type=int, default=0, help="number of last dispatches to average")
    ap.add_argument("--pmc-fetch", help="FETCH_SIZE counter CSV of the workload")
    ap.add_argument("--pmc-write", help="WRITE_SIZE counter CSV of the workload")
    ap.add_argument("--copy-fetch", help="FETCH_SIZE counter CSV of the copy calibration")
    ap.add_argument("--copy-write", help="WRITE_SIZE counter CSV of the copy calibration")
    ap.add_argument("--copy-bytes", type=int, default=0, help="bytes of the calibration copy")
    ap.add_argument("--json", help="write the traffic figures to this file (bench.py --traffic reads it)")
    a = ap.parse_args()
    if a.pmc_fetch:
        t = hbm_traffic(a.pmc_fetch, a.pmc_write, a.copy_fetch, a.copy_write, a.copy_bytes, a.kernel, a.last)
        print(json.dumps(t))
        if a.json:
            with open(a.json, "w") as f:
                json.dump(t, f)
    else:
        print(kernel_trace_summary(a.trace_csv, a.kernel, a.last))
