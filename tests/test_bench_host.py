"""bench.py's host-side plumbing: the kernel names a tuning launches and the
HBM traffic it reads from the committed rocprof summary."""
import os
import sys

from conftest import REPO

sys.path.insert(0, REPO)
import bench  # noqa: E402


def test_fused_kernel_names_follow_the_tuning_codes():
    assert bench.fused_kernel_names((0, 512, 8128)) == ("k_dir_x<512>", "k_dir_zt2<8, 128>")
    assert bench.fused_kernel_names((0, 1256, 18128)) == ("k_dir_xd<256>", "k_dir_zt2d<8, 128>")
    assert bench.fused_kernel_names((0, 1256, 18096)) == ("k_dir_xd<256>", "k_dir_zt1d<8, 96>")
    assert bench.fused_kernel_names((0, 512, 8096)) == ("k_dir_x<512>", "k_dir_zt1<8, 96>")
    assert bench.fused_kernel_names((0, 256, 4128)) == ("k_dir_x<256>", "k_dir_zt<4, 512>")
    assert bench.fused_kernel_names((0, 448, 8064)) == ("k_dir_x<448>", "k_dir_zt<8, 512>")
    assert bench.fused_kernel_names((0, 512, 16)) == ("k_dir_x<512>", "k_dir_z<16>")


def test_traffic_of_the_default_kernels_is_in_the_committed_summary():
    """The default shapes' kernels have a traffic line, and it lies between
    what a direction must move (read S, write the result: 64 B per cell,
    less the counters' few-percent spread) and the algorithmic 256 B per
    cell, at 2000 x 1000."""
    assert os.path.exists(bench.PROFILE_SUMMARY)
    cells = 2000 * 1000
    for k in bench.fused_kernel_names((0, 512, 8096)):
        t = bench.traffic_from_profile(k)
        assert t is not None, k
        assert 60 * cells <= t <= bench.BYTES_PER_CELL_LAUNCH * cells, (k, t)
    assert bench.traffic_from_profile("k_dir_x") is None  # exact names only


def test_valu_counts_of_the_default_kernels_are_in_the_committed_summary():
    """Each default kernel has its counted VALU instructions per launch, and
    the FP64 issue floor they imply lies below the kernel's traced average
    duration (a floor above it would mean the counter or the unit is wrong)."""
    for k in bench.fused_kernel_names((0, 512, 8096)):
        n = bench.valu_from_profile(k)
        assert n is not None and n > 1e6, k
        floor_us = n / bench.VALU_ISSUE_PEAK * 1e6
        assert 20.0 < floor_us < 60.0, (k, floor_us)
    assert bench.VALU_ISSUE_PEAK == 256 * 4 * 2.4e9 / 4


def test_roofline_units():
    assert bench.BYTES_PER_CELL_STEP == 2 * bench.BYTES_PER_CELL_LAUNCH == 512.0
    assert bench.FWD_ADJ_BYTES_PER_CELL_STEP == 512.0 + 512.0 + 768.0
    assert bench.HBM_PEAK_GBS == 8000.0
