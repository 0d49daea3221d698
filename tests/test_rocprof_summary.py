"""The rocprofv3 summary tool (ereb.rocprof_summary) on a synthetic kernel
trace: per-kernel means and totals, and one step's timeline between two
dispatches of a step's first kernel, with queue ids."""
from __future__ import annotations

import csv

import pytest

from ereb import rocprof_summary as rs

FIELDS = ["Kind", "Agent_Id", "Queue_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp",
          "Workgroup_Size_X", "Grid_Size_X", "Grid_Size_Y"]

# (queue, name, start ns, end ns): two steps of a first kernel A, with B on
# queue 2 overlapping C on queue 1
EVENTS = [
    (1, "A", 1_000, 3_000), (1, "C", 3_000, 7_000), (2, "B", 4_000, 9_000),
    (1, "A", 10_000, 12_500), (1, "C", 12_500, 16_500), (2, "B", 13_000, 18_000),
    (1, "A", 20_000, 22_000),
]


@pytest.fixture()
def trace(tmp_path):
    d = tmp_path / "trace"
    d.mkdir()
    with open(d / "run_kernel_trace.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, FIELDS)
        w.writeheader()
        for q, n, a, b in EVENTS:
            w.writerow({"Kind": "KERNEL_DISPATCH", "Agent_Id": 1, "Queue_Id": q,
                        "Kernel_Name": n, "Start_Timestamp": a, "End_Timestamp": b,
                        "Workgroup_Size_X": 256, "Grid_Size_X": 1024, "Grid_Size_Y": 1})
    return str(d)


def test_kernel_table(trace):
    tab = {name: (calls, mean, total) for name, _, calls, mean, total in rs.kernel_table(trace)}
    assert tab["A"] == (3, pytest.approx(6.5 / 3), pytest.approx(6.5))
    assert tab["B"] == (2, pytest.approx(5.0), pytest.approx(10.0))
    assert tab["C"] == (2, pytest.approx(4.0), pytest.approx(8.0))
    # sorted by total time
    assert [r[0] for r in rs.kernel_table(trace)] == ["B", "C", "A"]


def test_step_timeline(trace, tmp_path):
    tl = rs.step_timeline(trace, "A", 1)
    assert [(round(a, 3), round(b, 3), str(q), n) for a, b, q, n in tl] == [
        (0.0, 2.5, "1", "A"), (2.5, 6.5, "1", "C"), (3.0, 8.0, "2", "B")]
    out = tmp_path / "tl.md"
    rs.write_timeline(trace, "A", 0, "t", "cmd", str(out))
    text = out.read_text()
    assert "Span 8.0 µs, sum of the durations 11.0 µs" in text
    assert "| 3.0 | 8.0 | 5.0 | 2 | `B` |" in text
    with pytest.raises(ValueError):
        rs.step_timeline(trace, "A", 2)  # no dispatch of A after the last one
