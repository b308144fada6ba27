"""CPU test of the rocprof traffic summary (marie_vie.rocprof_summary) on a
synthetic profiler database: every pass of the operator is matched, for the
FP64 Stockham axis passes (axis_pass_ct) and the FP32 register-FFT ones
(axis_pass_rf); the bytes are calibrated by the plain copy; a database that
lacks a pass is refused instead of summed in part."""
import os
import sqlite3

import pytest

from marie_vie import rocprof_summary as rs

PASSES = {
    "fp64": ["void vie::axis_pass_ct<double, 2, 0, false, 112, 8>(vie::AxisArgs, HIP_vector_type<double, 2u> const*)",
             "void vie::axis_pass_ct<double, 0, 0, false, 96, 8>(vie::AxisArgs, HIP_vector_type<double, 2u> const*)",
             "void vie::x_fused_rf<double, true, false, false, 80>(vie::XArgs<double>, HIP_vector_type<double, 2u> const*)",
             "void vie::axis_pass_ct<double, 0, 0, true, 96, 8>(vie::AxisArgs, HIP_vector_type<double, 2u> const*)",
             "void vie::axis_pass_ct<double, 0, 3, true, 112, 8>(vie::AxisArgs, HIP_vector_type<double, 2u> const*)"],
    "fp32": ["void vie::axis_pass_rf<float, 2, 0, false, 112, 8>(vie::AxisArgs, HIP_vector_type<float, 2u> const*)",
             "void vie::axis_pass_rf<float, 0, 0, false, 96, 8>(vie::AxisArgs, HIP_vector_type<float, 2u> const*)",
             "void vie::x_fused_rf<float, true, false, false, 80>(vie::XArgs<float>, HIP_vector_type<float, 2u> const*)",
             "void vie::axis_pass_rf<float, 0, 0, true, 96, 8>(vie::AxisArgs, HIP_vector_type<float, 2u> const*)",
             "void vie::axis_pass_rf<float, 0, 3, true, 112, 8>(vie::AxisArgs, HIP_vector_type<float, 2u> const*)"],
}
COPY = "__amd_rocclr_copyBuffer"
CALIB = 512 << 20


def _db(path, counter, kernels, per_kib, copy_kib):
    os.makedirs(path)
    con = sqlite3.connect(os.path.join(path, "run_results.db"))
    con.execute("create table pmc_events (name text, dispatch_id int, counter_name text, counter_value real)")
    d = 0
    for k, v in zip(kernels, per_kib):
        for _ in range(3):
            d += 1
            con.execute("insert into pmc_events values (?,?,?,?)", (k, d, counter, v))
    con.execute("insert into pmc_events values (?,?,?,?)", (COPY, d + 1, counter, copy_kib))
    con.commit()
    con.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_traffic_sums_all_five_passes(tmp_path, prec):
    ks = PASSES[prec]
    fetch = [100.0, 200.0, 300.0, 400.0, 500.0]       # KiB per dispatch
    write = [10.0, 20.0, 30.0, 40.0, 50.0]
    # the copy reads FETCH_SIZE at half and WRITE_SIZE at full size
    _db(str(tmp_path / "f"), "FETCH_SIZE", ks, fetch, CALIB / 1024 / 2)
    _db(str(tmp_path / "w"), "WRITE_SIZE", ks, write, CALIB / 1024)
    text, total = rs.traffic(str(tmp_path / "f"), str(tmp_path / "w"), CALIB, prec)
    want = 1024.0 * (2.0 * sum(fetch) + sum(write))
    assert total == pytest.approx(want, rel=1e-9)
    assert "not found" not in text and text.count("\n") == 7


def test_traffic_refuses_a_missing_pass(tmp_path):
    ks = PASSES["fp32"][:2] + PASSES["fp32"][3:]     # no x pass
    _db(str(tmp_path / "f"), "FETCH_SIZE", ks, [1.0] * 4, 10.0)
    _db(str(tmp_path / "w"), "WRITE_SIZE", ks, [1.0] * 4, 10.0)
    with pytest.raises(SystemExit) as ei:
        rs.traffic(str(tmp_path / "f"), str(tmp_path / "w"), CALIB, "fp32")
    assert "x fused" in str(ei.value)


def test_counters_mean_per_dispatch(tmp_path):
    """counters: the mean per dispatch of every counter of two runs, for the
    matching kernel variant with the most dispatches."""
    a = os.path.join(tmp_path, "A")
    b = os.path.join(tmp_path, "B")
    x32 = PASSES["fp32"][2]
    _db(a, "SQ_WAVES", [x32, PASSES["fp32"][1]], [10.0, 20.0], 1.0)
    _db(b, "SQ_INSTS_VALU", [x32], [300.0], 1.0)
    text = rs.counters([a, b], [r"x_fused_rf<float", r"axis_pass_rf<float, 0, 0, false"])
    rows = {line.split()[0]: line.split()[1:] for line in text.splitlines()[1:]}
    assert float(rows["SQ_WAVES"][0]) == 10.0 and float(rows["SQ_WAVES"][1]) == 20.0
    assert float(rows["SQ_INSTS_VALU"][0]) == 300.0 and rows["SQ_INSTS_VALU"][1] == "nan"
    with pytest.raises(SystemExit):
        rs.counters([a], [r"no_such_kernel"])


def test_window_breaks_a_solver_loop_into_iterations(tmp_path):
    """Two warm-up launches, then three iterations of (op 5, step 2, gap 1):
    per iteration 8 units of span, 7 of kernels, 1 idle."""
    path = tmp_path / "w"
    os.makedirs(path)
    con = sqlite3.connect(os.path.join(path, "run_results.db"))
    con.execute("create table kernels (name text, start int, end int)")
    t = 0
    for it in range(5):
        con.execute("insert into kernels values (?,?,?)", ("op", t, t + 5_000_000))
        con.execute("insert into kernels values (?,?,?)", ("k_step", t + 5_000_000, t + 7_000_000))
        t += 8_000_000
    con.execute("insert into kernels values (?,?,?)", ("op", t, t + 5_000_000))
    con.commit()
    con.close()
    text = rs.window(str(path), "^op$", skip=2)
    first = text.splitlines()[0]
    assert "iterations 3" in first and "span 8.000 ms" in first and "idle 1.000 ms" in first
    assert "kernels 7.000 ms" in first
    body = {l.split()[0]: float(l.split()[1]) for l in text.splitlines()[1:]}
    assert body == {"op": 5.0, "k_step": 2.0}
