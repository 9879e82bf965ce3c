"""The command-line driver's host-side behaviour (src/main.cpp:41-83): usage,
--empty-prm, option checking and prm validation, all before any device work.
CPU only; the end-to-end runs are in test_gpu_driver_io.py."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
DRIVER = ROOT / "ctbb_icd_amd" / "ctbb_icd"

pytestmark = pytest.mark.skipif(not DRIVER.exists(), reason="driver not built")


def run(*args):
    r = subprocess.run([str(DRIVER), *map(str, args)], capture_output=True, text=True,
                       timeout=60)
    return r.returncode, r.stdout + r.stderr


def test_no_arguments_prints_usage():
    rc, out = run()
    assert rc == 0
    assert "usage: ctbb_icd" in out and "--empty-prm" in out


def test_unknown_option_prints_usage():
    rc, out = run("--bogus", "scan.yaml")
    assert rc == 0
    assert "usage: ctbb_icd" in out


def test_empty_prm_lists_the_keys():
    rc, out = run("--empty-prm")
    assert rc == 0
    for key in ("sinogram_path:", "fan_angle_increment:", "penalty:", "lambda:",
                "num_iterations:", "Zffs:", "Phiffs:", "wfbp_initialize:"):
        assert key in out, key


def test_prm_without_fan_angle_increment_is_rejected(tmp_path):
    """The reference's shipped sample prm lacks fan_angle_increment and fails
    validation (src/setup.cpp:129); the driver reports it and exits 1
    before touching a device."""
    from ctbb_icd_amd import configs
    txt = configs.prm_text("config0")
    txt = "\n".join(l for l in txt.splitlines() if not l.startswith("fan_angle_increment"))
    p = tmp_path / "no_fan.yaml"
    p.write_text(txt + "\n")
    rc, out = run(p)
    assert rc == 1
    assert "fan_angle_increment not properly set in configuration" in out
    assert "No GPU" not in out
