"""Host logic of the HBM-traffic summary (workloads.pmc_traffic): counter
scaling by the calibration copy and one entry per kernel instance."""
import json
import sqlite3

from workloads import pmc_traffic


def _db(path, counter, rows):
    con = sqlite3.connect(path)
    con.execute("create table counters_collection (kernel_name text, counter_name text, value real)")
    con.executemany("insert into counters_collection values (?, ?, ?)", [(k, counter, v) for k, v in rows])
    con.commit()
    con.close()


def test_scaling_and_template_instances(tmp_path):
    copy = "__amd_rocclr_copyBuffer"
    fwd = "void mpol::k_fft_y_xy<float, 0, 256, false>(mpol::ColArgs, mpol::ScaleParams, mpol::cplx<float>*)"
    inv = "void mpol::k_fft_y_xy<float, 0, 256, true>(mpol::ColArgs, mpol::ScaleParams, mpol::cplx<float>*)"
    one = "void mpol::k_real_space<float>(mpol::RealArgs<float>)"
    # the fetch counter reads 160 of the 256 MiB a copy moves, the write
    # counter all of it; values are KiB per dispatch
    _db(tmp_path / "f.db", "FETCH_SIZE",
        [(copy, 160 * 1024), (copy, 4.0), (fwd, 100.0), (fwd, 300.0), (inv, 50.0), (one, 10.0)])
    _db(tmp_path / "w.db", "WRITE_SIZE",
        [(copy, 256 * 1024), (fwd, 20.0), (fwd, 20.0), (inv, 70.0), (one, 1.0)])
    out = tmp_path / "t.json"
    pmc_traffic.main([str(tmp_path / "f.db"), str(tmp_path / "w.db"), "256", str(out)])
    res = json.loads(out.read_text())
    assert abs(res["fetch_scale"] - 1.6) < 1e-12 and abs(res["write_scale"] - 1.0) < 1e-12
    k = res["kernels"]
    assert set(k) == {"k_fft_y_xy<float, 0, 256, false>", "k_fft_y_xy<float, 0, 256, true>", "k_real_space"}
    f = k["k_fft_y_xy<float, 0, 256, false>"]
    assert f["launches"] == 2
    assert f["fetch_bytes"] == round(1024 * 1.6 * 200.0) and f["write_bytes"] == 1024 * 20
    assert k["k_fft_y_xy<float, 0, 256, true>"]["bytes_per_launch"] == round(1024 * 1.6 * 50) + 1024 * 70
    assert k["k_real_space"]["bytes_per_launch"] == round(1024 * 1.6 * 10) + 1024
