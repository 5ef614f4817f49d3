"""bench.py's HBM traffic leg, on the CPU with a stand-in profiler.

The leg runs `rocprofv3 --pmc <counter> ... -- python bench.py <args>` four
times (FETCH_SIZE and WRITE_SIZE of the correlation launch, then of a 1 GiB
device copy) and scales each counter by the copy's bytes over its reading.
Here a small script named rocprofv3 first on PATH writes the counter CSV
that rocprofv3 writes (one row per dispatch and counter, values in KiB), so
the command lines, the CSV parsing, the choice of the largest matching
dispatch and the scaling are checked without a GPU.
"""
import os
import stat
import sys
import textwrap

import pytest

import bench

FAKE = textwrap.dedent('''\
    #!{python}
    import os, sys
    args = sys.argv[1:]
    counter = args[args.index("--pmc") + 1]
    out = os.path.join(args[args.index("-d") + 1], "host", "1234")
    child = args[args.index("--") + 1:]
    assert child[0] == sys.executable or child[0].endswith("python3") or "python" in child[0]
    assert child[1].endswith("bench.py")
    os.makedirs(out, exist_ok=True)
    copy = "--copy-calibration" in child
    rows = []
    if copy:
        assert child[child.index("--copy-calibration") + 1] == "1024"
        kib = {{"FETCH_SIZE": 524288.0, "WRITE_SIZE": 1048576.0}}[counter]
        rows.append(("__amd_rocclr_copyBuffer", kib))
        rows.append(("__amd_rocclr_copyBuffer", kib / 1024))
    else:
        assert "--no-traffic" in child and "--steps" in child
        kib = {{"FETCH_SIZE": 1000.0, "WRITE_SIZE": 500.0}}[counter]
        rows.append(("void fmf::corr_mfma32_kernel<1, 4, 2, 2, 3, 1>(...)", kib))
        rows.append(("void fmf::window_energy_union_kernel<0, 1>(...)", 10 * kib))
    with open(os.path.join(out, "run_counter_collection.csv"), "w") as f:
        f.write('"Dispatch_Id","Kernel_Name","Counter_Name","Counter_Value"\\n')
        for i, (name, v) in enumerate(rows):
            f.write('%d,"%s","%s",%f\\n' % (i + 1, name, counter, v))
''')


@pytest.fixture
def fake_rocprof(tmp_path, monkeypatch):
    path = tmp_path / "rocprofv3"
    path.write_text(FAKE.format(python=sys.executable))
    path.chmod(path.stat().st_mode | stat.S_IEXEC | stat.S_IXGRP | stat.S_IXOTH)
    monkeypatch.setenv("PATH", str(tmp_path) + os.pathsep + os.environ.get("PATH", ""))
    return path


def test_traffic_leg_scales_counters_by_the_copy(fake_rocprof):
    child = ["--workload", "batched", "--step", "1", "--kernel", "bf16x3", "--steps", "1",
             "--warmup", "0", "--no-traffic"]
    r = bench.measure_traffic(child, "corr_mfma", timeout=60)
    # FETCH_SIZE of the 1 GiB copy reads 524288 KiB: x2; WRITE_SIZE 1048576 KiB: x1
    assert r["fetch_scale"] == pytest.approx(2.0)
    assert r["write_scale"] == pytest.approx(1.0)
    # the correlation dispatch only, not the larger energy dispatch
    assert r["fetch_KiB"] == 1000.0 and r["write_KiB"] == 500.0
    assert r["bytes"] == pytest.approx(1000 * 1024 * 2 + 500 * 1024 * 1)


def test_traffic_leg_reports_a_missing_kernel(fake_rocprof):
    child = ["--workload", "batched", "--steps", "1", "--no-traffic"]
    with pytest.raises(RuntimeError, match="no FETCH_SIZE dispatch"):
        bench.measure_traffic(child, "corr_fp32", timeout=60)
    # as a leg of the line: an error entry, and the recorded figure stays usable
    leg = bench.optional_leg("traffic", lambda: bench.measure_traffic(child, "corr_fp32",
                                                                        timeout=60))
    assert "error" in leg
