"""Host-side behaviour of bench.py that needs no GPU: a multi-GPU request is
refused with a message before anything touches a device, and the committed
rocprofv3 summary that feeds the bench line's `traffic` and `valu_issue`
fields is read for its own workload and for no other."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

FACES_4096 = 11249862  # faces of the 4096-chunk workload the summary was measured on


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert r.stdout == ""  # no JSON line
    assert "one GPU" in r.stderr


def test_traffic_summary_is_read_for_its_workload_only():
    fill = bench.measured_traffic("fill_kernel<32>", FACES_4096)
    emit = bench.measured_traffic("emit_kernel<32>", FACES_4096)
    insts = bench.measured_counter("fill_kernel<32>", FACES_4096, "valu_insts")
    assert fill is not None and emit is not None and insts is not None
    # HBM bytes within 2% of the algorithmic bytes of one launch (fill: ids +
    # occupancy + planes + meta; emit: 104 B/face + rows and planes read);
    # the rest is each workgroup's chunk descriptor and the read-only tables
    n, n3, n2 = 4096, 32 ** 3, 32 ** 2
    fill_alg = n * (n3 + n3 // 8 + 4 * n2 // 8 + 4)
    emit_alg = n * (n3 // 8 + 4 * n2 // 8 + 4 + 8) + 104 * FACES_4096
    assert abs(fill / fill_alg - 1) < 0.02
    assert abs(emit / emit_alg - 1) < 0.02
    # a different workload (face count) gets no figure
    assert bench.measured_traffic("fill_kernel<32>", FACES_4096 + 1) is None
    assert bench.measured_traffic("no_such_kernel", FACES_4096) is None
