"""Host logic of bench.py: the one-GPU contract, the committed traffic figure
of the headline workload, and the census flop count."""
from __future__ import annotations

import argparse
import subprocess
import sys
from pathlib import Path
from types import SimpleNamespace

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_traffic_figure_only_for_the_profiled_workload():
    args = argparse.Namespace(records=192, rays_log2=26, theta=900, phi=1800)
    t = bench.drop_traffic(args)
    assert t is not None and t > 1e11
    args.records = 12
    assert bench.drop_traffic(args) is None


def test_census_counts_each_event_kind():
    c = bench.CENSUS["drop"]
    st = SimpleNamespace(status=[0] * 8 + [10], rays=110, internal_hits=300, samples=250)
    want = 10 * c["miss"] + 100 * c["entry"] + 300 * c["hit"] + 250 * c["emit"]
    assert bench.census_flops([st], ["drop"]) == want
