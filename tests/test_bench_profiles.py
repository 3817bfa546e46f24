"""The committed rocprofv3 summaries that bench.py reads (HBM traffic and
VALU instruction counts per launch) describe the shipped launch shape, so
the benchmark line carries measured figures rather than nulls.  CPU only."""
import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def shipped_tuning():
    table = json.load(open(os.path.join(ROOT, "yoricktrtt_amd", "tuning", "tuning.json")))
    key = next(k for k in table if "cone:256x256x256x1:384x512:360" in k)
    return table[key]["tuning"]


def test_profiles_match_shipped_tuning():
    tu = shipped_tuning()
    for kernel in ("axial::fwd", "axial::bwd"):
        traffic = bench.measured_traffic(tu, kernel)
        assert traffic is not None and traffic > 0
        issue = bench.measured_issue(tu, kernel, 10.0)
        assert issue is not None
        assert issue["peak"] == pytest.approx(256 * 4 * 2.4e9 / 4)
        assert issue["achieved"] == pytest.approx(issue["insts_per_launch"] / 10.0e-3, rel=1e-3)


def test_profiles_ignored_for_another_launch_shape():
    tu = dict(shipped_tuning())
    tu["fwd_window_bytes"] += 4096
    assert bench.measured_traffic(tu, "axial::fwd") is None
    assert bench.measured_issue(tu, "axial::fwd", 10.0) is None
    assert bench.measured_issue(None, "axial::bwd", 10.0) is None


def test_forward_traffic_within_the_per_view_bytes():
    vals = bench.profile_values(bench.TRAFFIC_FILE, shipped_tuning())
    # the windows go to the projections with atomics (no scratch round trip):
    # fetched plus written bytes of the forward stay below one volume read
    # and one projection write per view
    assert vals["fwd_launches"] == 1
    assert 0 < vals["fwd_traffic_bytes_per_launch"] <= vals["fwd_algorithmic_bytes_per_launch"]


def test_footprint_weights_match_the_fixture():
    # bench.py's flop count is the frozen oracle count of the same workload
    doc = json.load(open(os.path.join(HERE, "golden", "footprint_count_cone256.json")))
    assert bench.FOOTPRINT_WEIGHTS == doc["total"]
    assert (bench.NX, bench.NY, bench.NZ, bench.NU, bench.NV, bench.NVIEWS) == (256, 256, 256, 384,
                                                                                 512, 360)
    assert (bench.RSC, bench.RSD) == (1000.0, 1500.0)
    import importlib.util

    import numpy as np

    spec = importlib.util.spec_from_file_location(
        "make_footprint_count", os.path.join(HERE, "golden", "make_footprint_count.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    a, b = bench.geometry(), mk.bench_geometry()
    assert np.array_equal(a.xos, b.xos) and np.array_equal(a.xsd, b.xsd)
