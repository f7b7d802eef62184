"""The bench line's bookkeeping, without a GPU: every stage bench.py times maps to a kernel
whose HBM traffic the committed counter summary of the headline workload holds, so a
renamed kernel cannot silently drop `traffic` from the roofline object; the summary's
calibration is the copy kernel's, read + write equal to twice the bytes it copies."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_every_stage_has_profiled_traffic():
    grid = (512, 1024, 1024)
    for stage in bench.STAGE_KERNEL:
        traffic, src = bench.profiled_traffic(grid, stage)
        assert traffic is not None and traffic > 0, stage
        assert src == os.path.join("profiles", "step_512x1024x1024_hbm_traffic.md")


def test_copy_kernel_calibration_in_summary():
    path = os.path.join(ROOT, "profiles", "step_512x1024x1024_hbm_traffic.md")
    rows = {}
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch "):
            _, name, value = line.split()
            rows[name] = float(value)
    # the copy probe moves 4 GiB each way per dispatch; scaled counters must say exactly that
    assert abs(rows["copy"] - 2 * (4 << 30)) <= 1e-6 * rows["copy"]
    # the dominant kernel's traffic is at least its algorithmic bytes (14.33 words per cell)
    assert rows["predict_w"] >= 14.33 * 8 * 512 * 1024 * 1024
