"""The bench line's counter traffic comes from committed rocprofv3 summaries;
a row carries it only when it ran the configuration the summary measured."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_every_profiled_row_has_a_traffic_figure():
    for row, (name, cfg) in bench.PROFILE_OF.items():
        assert os.path.exists(os.path.join(ROOT, "profiles", name)), name
        t = bench.traffic_from_profiles(row, dict(cfg, xcd=1, nt=0, wgpc=0))
        assert t is not None and t > 1e9, row


def test_another_configuration_gets_no_traffic_figure():
    for row, (_, cfg) in bench.PROFILE_OF.items():
        for key in ("kind", "threads", "pf"):
            other = dict(cfg)
            other[key] = cfg[key] + 1
            assert bench.traffic_from_profiles(row, other) is None, (row, key)
