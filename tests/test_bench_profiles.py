"""Host logic of bench.py's rocprofv3 summaries: stage attribution by launch order and the per-batch
figures of --summarize-batch-profiles, on small synthetic CSV directories (no GPU)."""
import csv
import json
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402

FRAME = [  # (kernel, duration ns, fetch KiB, write KiB): one batch frame in launch order
    ("void gs::(anonymous namespace)::k_project_groups<3>(gs::ViewDev const*)", 600, 100.0, 50.0),
    ("void gs::(anonymous namespace)::k_rts_up<8u>(unsigned int const*)", 30, 10.0, 1.0),
    ("gs::(anonymous namespace)::k_rts_scan(unsigned int*)", 10, 1.0, 1.0),
    ("void gs::(anonymous namespace)::k_rts_down<8u>(unsigned int const*)", 140, 20.0, 20.0),
    ("gs::(anonymous namespace)::k_tile_counts(gs::ViewDev const*)", 400, 200.0, 5.0),
    ("void gs::(anonymous namespace)::k_rts_down<9u>(unsigned int const*)", 500, 40.0, 40.0),
    ("gs::(anonymous namespace)::k_raster_order(uint2 const*)", 5, 0.5, 0.5),
    ("void gs::(anonymous namespace)::k_raster_rows2<false>(gs::ViewDev const*)", 2000, 300.0, 100.0),
    ("__amd_rocclr_copyBuffer", 4, 0.0, 0.0),
]


def _write(d: pathlib.Path, name: str, header, rows):
    d.mkdir(parents=True, exist_ok=True)
    with open(d / name, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def _dirs(tmp_path, frames=2):
    trace, fetch, write = tmp_path / "trace", tmp_path / "fetch", tmp_path / "write"
    tr, fe, wr = [], [], []
    t, i = 1000, 0
    for _ in range(frames):
        for name, ns, fk, wk in FRAME:
            i += 1
            tr.append([i, name, t, t + ns])
            fe.append([i, name, fk])
            wr.append([i, name, wk])
            t += ns + 1
    _write(trace, "b_kernel_trace.csv", ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"], tr)
    _write(fetch, "b_counter_collection.csv", ["Dispatch_Id", "Kernel_Name", "Counter_Value"], fe)
    _write(write, "b_counter_collection.csv", ["Dispatch_Id", "Kernel_Name", "Counter_Value"], wr)
    return trace, fetch, write


def test_stage_sequence_launch_order():
    rows = [{"Dispatch_Id": str(i), "Kernel_Name": name} for i, (name, *_rest) in enumerate(FRAME)]
    rows.append({"Dispatch_Id": "99", "Kernel_Name": "void gs::k_raster<true, 1>(gs::ViewDev const*)"})
    stages = [s for s, _ in bench._stage_sequence(rows)]
    assert stages == ["project", "sort", "sort", "sort", "bin", "bin", "bin", "raster", None, "counting"]


def test_stage_sequence_one_row_raster_closes_frame():
    rows = [{"Dispatch_Id": "1", "Kernel_Name": "void gs::k_project<3>(x)"},
            {"Dispatch_Id": "2", "Kernel_Name": "void gs::k_os_hist<8u>(x)"},
            {"Dispatch_Id": "3", "Kernel_Name": "gs::k_tile_counts(x)"},
            {"Dispatch_Id": "4", "Kernel_Name": "void gs::k_raster<false, 1>(x)"},
            {"Dispatch_Id": "5", "Kernel_Name": "void gs::k_rts_up<8u>(x)"}]
    assert [s for s, _ in bench._stage_sequence(rows)] == ["project", "sort", "bin", "raster", None]


def test_summarize_batch_profiles(tmp_path):
    trace, fetch, write = _dirs(tmp_path, frames=2)
    out = tmp_path / "out"
    bench.summarize_batch_profiles("batched-64v", trace, fetch, write, out, steps=1, warmup=1)
    text = (out / "kernels_batched-64v.md").read_text()
    scale = json.loads(bench.TRAFFIC_FILE.read_text())
    ff, wf = float(scale["fetch_scale"]), float(scale["write_scale"])
    rows = {}
    for line in text.splitlines():
        cells = [c.strip() for c in line.strip("|").split("|")]
        if len(cells) >= 4 and cells[0] and not cells[0].startswith("-"):
            rows[cells[0]] = cells[1:]
    # per kernel: calls per batch, avg us, us per batch, %, fetch MB, write MB (per batch)
    tc = rows["k_tile_counts"]
    assert float(tc[0]) == pytest.approx(1.0)
    assert float(tc[1]) == pytest.approx(0.4, abs=0.01)
    assert float(tc[4]) == pytest.approx(200.0 * 1024 * ff / 1e6, abs=0.05)
    assert float(tc[5]) == pytest.approx(5.0 * 1024 * wf / 1e6, abs=0.05)
    # per stage (us per batch, fetch MB, write MB): the rows2 raster is the raster stage
    assert float(rows["k_raster"][0]) == pytest.approx(2.0, abs=0.05)
    assert float(rows["depth sort"][0]) == pytest.approx(0.2, abs=0.05)
    assert float(rows["bin(pair sort+emit)"][0]) == pytest.approx(0.9, abs=0.05)
    assert float(rows["bin(pair sort+emit)"][1]) == pytest.approx(240.5 * 1024 * ff / 1e6, abs=0.05)
    assert float(rows["k_project"][2]) == pytest.approx(50.0 * 1024 * wf / 1e6, abs=0.05)


def test_batch_workload_names():
    with pytest.raises(SystemExit):
        bench.batch_views("no-such-workload")
