"""bench.py's command line without a device: more than one GPU is refused
with a message, before anything touches a GPU."""
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_cpu_matrix_build_leg_runs_on_a_view_sample(make_scan):
    """The single-threaded matrix-build baseline times whole views of the
    ray-driven build and reports rays/s over the views it traced."""
    sys.path.insert(0, str(ROOT))
    import bench
    g = make_scan("tiny_ffs").geom
    r = bench.cpu_ssm_build(g, 0.05)
    assert r["unit"] == "rays/s" and r["cores"] == 1 and r["kind"] == "port"
    assert r["value"] > 0
    nv = int(r["sample"].split()[0])
    assert 1 <= nv <= g.n_views_sm
    assert f"{nv * g.n_channels * g.n_rows} rays" in r["sample"]
