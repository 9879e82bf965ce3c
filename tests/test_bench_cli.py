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


def test_cpu_icd_legs_in_ram_and_streamed(make_scan, orc):
    """The CPU ICD baseline on a band of rows, matrix in RAM and read from a
    matrix.bin file of the band every iteration (the device matrix's column
    download stood in for by the oracle's matrix)."""
    sys.path.insert(0, str(ROOT))
    import bench
    s = make_scan("tiny_ffs")
    cp, idx, val = orc.ssm_build(s.geom)

    class Columns:
        def download_cols(self, c0, c1):
            return cp[c0:c1 + 1] - cp[c0], idx[cp[c0]:cp[c1]], val[cp[c0]:cp[c1]]

    raw, w = s.synthesize(use_gpu=False)
    r = bench.cpu_baseline(s, Columns(), raw, w, 0.05, 2, streamed=True)
    assert r["unit"] == "voxel updates/s" and r["cores"] == 2 and r["value"] > 0
    assert r["streamed"]["value"] > 0
    upd = int(r["sample"].split("first iteration from zero, ")[1].split()[0])
    assert f"{upd} updates" in r["streamed"]["sample"]
