"""CPU checks of bench.py's command line and workload helpers (no GPU needed): a request
for more than one GPU is refused before any device work, every workload is offered, and the
batching/serving job mix follows its stated distribution (resolutions from the four sizes,
three generators, spp 16..256)."""
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    assert "one GPU only" in p.stderr
    assert p.stdout == ""


def test_every_workload_is_offered():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0
    for w in ("cornell_1080p", "batch256", "serve", "book_640", "sdf_720p"):
        assert w in p.stdout


def test_job_mix_follows_its_distribution():
    jobs = bench.draw_jobs(4000, 7)
    sizes = {(j["W"], j["H"]) for j in jobs}
    assert sizes == set(bench.RESOLUTIONS)
    gens = np.bincount([j["gen"] for j in jobs], minlength=3)
    assert gens.min() > 1200  # uniform over the three generators
    spp = np.array([j["spp"] for j in jobs])
    assert spp.min() >= 16 and spp.max() <= 256 and abs(spp.mean() - 136) < 5
    assert bench.draw_jobs(10, 7) == bench.draw_jobs(10, 7)  # seeded


def test_percentile_matches_numpy():
    xs = [5.0, 1.0, 9.0, 3.0, 7.0]
    assert bench.percentile(xs, 50) == 5.0
    assert bench.percentile(xs, 99) == float(np.percentile(np.array(xs), 99))
    assert bench.percentile([], 99) is None
