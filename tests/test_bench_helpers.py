"""bench.py's arithmetic and its reading of the committed profiles (no GPU)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_profile_figures_are_read_from_the_committed_summaries():
    # the traversal summary (.txt) and the cache kernels' summary (.md) both hold machine lines
    for name, kernel in [("valu_inst_per_ray", "k_extend"), ("valu_lane_util_pct", "k_extend"),
                         ("hbm_bytes_per_ray", "k_shadow"), ("l2_hit_pct", "k_extend"),
                         ("mfma_tflops", "k_mlp_train"), ("mfma_busy_pct", "k_mlp_infer")]:
        v = bench.profile_figure(name, kernel)
        assert v is not None and v > 0, (name, kernel)
    assert 0 < bench.profile_figure("valu_lane_util_pct", "k_extend") <= 100
    assert bench.profile_figure("no_such_figure", "k_extend") is None


def test_stage_roofline_against_the_valu_issue_peak():
    """achieved = VALU wave-instructions per ray x rays per launch / mean launch time; the peak
    is 1024 SIMDs x 2.4 GHz / 4 cycles; the HBM figure is the profile's bytes per ray x rays
    over the same time against 8 TB/s."""
    assert bench.PEAK_VALU_GINST == pytest.approx(614.4)
    d = {"launches": 10, "ms": 20.0, "rays": 1.0e8, "bytes": 2.0e11}
    r = bench.stage_roofline("k_extend", d)
    per_ray = bench.profile_figure("valu_inst_per_ray", "k_extend")
    hbm_ray = bench.profile_figure("hbm_bytes_per_ray", "k_extend")
    avg_s, rays = 20.0e-3 / 10, 1.0e8 / 10
    assert r["bound"] == "valu" and r["peak"] == pytest.approx(614.4)
    assert r["achieved"] == pytest.approx(per_ray * rays / avg_s / 1e9, rel=1e-3)
    assert r["frac"] == pytest.approx(r["achieved"] / 614.4, rel=1e-3)
    assert r["traffic"] == pytest.approx(hbm_ray * rays, rel=1e-6)
    assert r["hbm"]["frac"] == pytest.approx(hbm_ray * rays / avg_s / 1e9 / 8000.0, rel=1e-3)
    assert r["demand_bytes_per_launch"] == pytest.approx(2.0e10)
    assert r["avg_launch_ms"] == pytest.approx(2.0)


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr + p.stdout
