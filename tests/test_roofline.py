"""The algorithmic byte counts behind the bench's roofline figures
(scisim_amd/roofline.py), checked against the per-unit figures derived field by field
from the record layouts: a dropped or doubled term changes one of these numbers."""

from scisim_amd import roofline as rl


def test_sweep_bytes_per_contact():
    # 2D FP32: record 32 (ids 8, n 8, W 4, forcing 8, mu 4), impulse read + write 2*2*4,
    # both bodies' 2-component velocity deltas read + write 2*2*2*4
    assert rl.gs_sweep_bytes(2, True) == 32 + 16 + 32 == 80
    # 3D FP32: record 64 (ids 8, n and s 24, arms 8, W_n W_t 8, forcing 12, mu 4), impulse
    # 2*3*4, both bodies' 6 dof 2*2*6*4
    assert rl.gs_sweep_bytes(3, True) == 64 + 24 + 96 == 184
    # an FP64 sweep counts twice the FP32 figure
    assert rl.gs_sweep_bytes(2, False) == 160 and rl.gs_sweep_bytes(3, False) == 368


def test_residual_bytes_per_contact():
    # record, impulse and both bodies read in FP64: 32 + 8*2 + 2*8*2 and 64 + 8*3 + 2*8*6
    assert rl.gs_eval_bytes(2, True) == rl.gs_eval_bytes(2, False) == 80
    assert rl.gs_eval_bytes(3, True) == rl.gs_eval_bytes(3, False) == 184


def test_launch_bytes_add_up():
    C, S = 1000, 7
    assert rl.gs_launch_bytes(2, C, S, S + 1, True) == C * (S * rl.gs_sweep_bytes(2, True) +
                                                          (S + 1) * rl.gs_eval_bytes(2, True))
    assert rl.gs_launch_bytes(3, 0, S, S + 1, False) == 0


def test_pairs_bytes():
    n, G, E, P = 1000, 500, 4000, 300
    per_body = 4 + 16 + 40 + 4
    assert rl.pairs_bytes_per_body(2) == per_body
    assert rl.pairs_bytes_per_body(3) == 4 + 24 + 32 + 4
    assert rl.pairs_launch_bytes(2, n, G, E, P, False) == n * per_body + 4 * (G + 1) + 4 * E
    assert rl.pairs_launch_bytes(2, n, G, E, P, True) - rl.pairs_launch_bytes(2, n, G, E, P, False) == 8 * P


def test_broadphase_bytes():
    # per body: predictor 52/72, three 16-byte radix passes 48, cell ranges 16/12,
    # neighbour sweep 48/68, ordered output 8; per pair 8
    assert rl.broadphase_bytes(2, 1, 0) == 52 + 48 + 16 + 48 + 8 == 172
    assert rl.broadphase_bytes(3, 1, 0) == 72 + 48 + 12 + 68 + 8 == 208
    assert rl.broadphase_bytes(2, 10, 3) == 1720 + 24


def test_kernel_trace_summary_averages_the_last_dispatches(tmp_path):
    import csv
    p = tmp_path / "trace.csv"
    with open(p, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kind", "Agent_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        t = 0
        for k in range(6):
            w.writerow(["KERNEL_DISPATCH", 1, "void scx::k_gs<3, float>(scx::GSArgs<float>)", t, t + 1000 * (k + 1)])
            t += 10000
            w.writerow(["KERNEL_DISPATCH", 1, "k_pairs", t, t + 500])
            t += 1000
    out = rl.kernel_trace_summary(str(p), "k_gs", 2)
    assert "| `k_gs<3, float>` | 6 | 3.5 | 1.0 | 6.0 | 21.0 |" in out
    assert "Last 2 dispatches of `k_gs*` (the timed steps): mean 5.5 us" in out


def test_hbm_traffic_from_counter_csvs(tmp_path):
    """Per-launch HBM bytes from FETCH_SIZE / WRITE_SIZE runs (KiB per dispatch), each scaled
    by what the counter reads for a plain copy of known size; the last dispatches only."""
    from scisim_amd import roofline as rl
    head = "Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value\n"

    def write(name, rows):
        p = tmp_path / name
        p.write_text(head + "".join('%d,"%s",%s,%s\n' % r for r in rows))
        return str(p)

    # the copy of 1 MiB reads 512 KiB by FETCH_SIZE (factor 2) and writes 1024 KiB (factor 1)
    cf = write("cf.csv", [(1, "__amd_rocclr_copyBuffer", "FETCH_SIZE", 512), (2, "__amd_rocclr_copyBuffer",
                                                                             "FETCH_SIZE", 512)])
    cw = write("cw.csv", [(1, "__amd_rocclr_copyBuffer", "WRITE_SIZE", 1024)])
    # k_gs: an early dispatch (dropped by last=2) and two timed ones, one split over two rows
    kf = write("kf.csv", [(5, "void scx::k_gs<3, float, true>(scx::GSArgs<float>)", "FETCH_SIZE", 999),
                          (7, "void scx::k_gs<3, float, true>(scx::GSArgs<float>)", "FETCH_SIZE", 100),
                          (7, "void scx::k_gs<3, float, true>(scx::GSArgs<float>)", "FETCH_SIZE", 100),
                          (8, "void scx::k_gs<3, double, true>(scx::GSArgs<double>)", "FETCH_SIZE", 400),
                          (9, "void scx::k_colour(unsigned int)", "FETCH_SIZE", 7)])
    kw = write("kw.csv", [(7, "void scx::k_gs<3, float, true>(scx::GSArgs<float>)", "WRITE_SIZE", 50),
                          (8, "void scx::k_gs<3, double, true>(scx::GSArgs<double>)", "WRITE_SIZE", 150)])
    t = rl.hbm_traffic(kf, kw, cf, cw, 1 << 20, "k_gs", last=2)
    assert t["fetch_factor"] == 2.0 and t["write_factor"] == 1.0
    assert t["dispatches"] == [2, 2]
    assert t["fetch_bytes"] == 1024.0 * (200 + 400) / 2 * 2.0
    assert t["write_bytes"] == 1024.0 * (50 + 150) / 2
    assert t["bytes_per_launch"] == t["fetch_bytes"] + t["write_bytes"]
