"""prm interface: parse semantics, validation, derived geometry, slice table.

Derived values are checked against numbers stated independently of the code
(the per-config size table worked out from src/setup.cpp:183-237 by hand) and
against closed-form invariants of the slice table (src/setup.cpp:349-409)."""
import math
import subprocess
import sys
import textwrap
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def run_configure(path):
    code = textwrap.dedent(f"""
        import sys; sys.path.insert(0, {str(ROOT)!r})
        import ctbb_icd_amd as ct, ctypes
        g = ct.Geom(); rp = ct.RunParams()
        rc = ct.lib().ctbb_configure({str(path)!r}.encode(), ctypes.byref(g), ctypes.byref(rp),
                                     None, None, None, None)
        print("RC", rc, g.nz, g.n_views_sm)
    """)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
    return r.stdout


@pytest.mark.parametrize("name,nsm,vps,nz,nxy,rows,readings", [
    ("config0", 700, 36, 32, 128, 16, 2304),
    ("config1", 3000, 36, 96, 512, 64, 6912),
    ("config3", 3000, 36, 32, 256, 64, 4608),
    ("config4", 3000, 36, 256, 512, 64, 13824),
    ("tiny", 100, 12, 8, 24, 8, 288),
    ("tiny_ffs", 100, 12, 8, 24, 8, 288),
    # z flying focal spot: 96 views/turn, beam 16 x 0.22 cm = 3.52 cm, at the
    # ROI edge 3.52 * (59.5 + 2.4) / 108.56 = 2.007 -> N_sm = int(100 * 2.007)
    # = 200; 8 projection rows; slices 0.12 / (0.96 / 96) = 12 views
    ("tiny_zffs", 200, 12, 5, 24, 8, 288),
    # both spots: 192 views/turn -> int(200 * 2.007) = 401, raised to 408
    # (even, N_sm/2 divisible by n_ffs = 4); slices 0.12 / 0.005 = 24 views
    ("tiny_zphiffs", 408, 24, 5, 24, 8, 576),
])
def test_derived_sizes(make_scan, name, nsm, vps, nz, nxy, rows, readings):
    s = make_scan(name)
    g = s.geom
    assert g.n_views_sm == nsm
    assert s.run.views_per_slice == vps
    assert (g.nx, g.ny, g.nz) == (nxy, nxy, nz)
    assert g.n_rows == rows and g.readings == readings
    nffs = 2 ** (g.zffs + g.phiffs)
    assert (g.n_views_sm // 2) % nffs == 0 and g.n_views_sm % 2 == 0
    assert g.center_voxel_x == (g.nx - 1) / 2 and g.center_voxel_z == 0.0
    assert math.isclose(g.voxel_size_z, vps * g.tube_z_increment, rel_tol=1e-12)
    assert g.table_direction == -1


def test_slice_table_invariants(make_scan):
    for name in ["config0", "config1", "tiny_ffs"]:
        s = make_scan(name)
        g = s.geom
        nffs = 2 ** (g.zffs + g.phiffs)
        si = s.slice_index
        assert np.all(si % nffs == 0)
        # central view: first (multiple of nffs) with table position <= location
        tp = s.table_positions
        for k in range(g.nz):
            assert tp[si[k]] <= s.slice_locations[k]
            assert si[k] == 0 or tp[si[k] - nffs] > s.slice_locations[k]
        # FileType-0 angles: (360/vpt) i in float, mod 360
        vpt = s.run.num_views_per_turn
        i = np.arange(g.readings)
        ang = np.fmod((np.float32(360.0) / np.float32(vpt) * i.astype(np.float32)).astype(np.float64), 360.0)
        assert np.array_equal(ang, s.tube_angles)
        # every slice's full footprint lies inside the scan
        assert si.min() - g.n_views_sm // 2 >= 0
        assert si.max() + g.n_views_sm // 2 <= g.readings


def test_range_check_rejects_two_rotations(prm_dir, ct):
    """Config 0 as literally stated (2 rotations) fails the reference's own
    allowed-z-range check; 4 rotations pass."""
    from ctbb_icd_amd import configs
    p = configs.write_prm("config0", prm_dir / "c0_2rot.yaml", Readings=1152)
    out = run_configure(p)
    assert "is invalid" in out and "Allowed reconstruction range" in out
    assert "RC -1" in out


def test_missing_fan_angle_increment_fails(prm_dir):
    """A prm with no fan_angle_increment (like the reference's shipped sample)
    fails validation: the key is required (src/setup.cpp:129)."""
    from ctbb_icd_amd import configs
    txt = configs.prm_text("config0")
    txt = "\n".join(l for l in txt.splitlines() if not l.startswith("fan_angle_increment"))
    p = prm_dir / "no_fan.yaml"
    p.write_text(txt + "\n")
    out = run_configure(p)
    assert "fan_angle_increment not properly set in configuration" in out
    assert "Missing some required parameters" in out
    assert "RC -1" in out


def test_found_not_found_printout(prm_dir):
    from ctbb_icd_amd import configs
    p = configs.write_prm("tiny", prm_dir / "tiny_print.yaml")
    out = run_configure(p)
    lines = out.splitlines()
    assert lines[0].startswith("Config File: ")
    nx = [l for l in lines if l.startswith("nx:")][0]
    assert nx == "nx:" + " " * (1 + 35 - 3) + "FOUND "
    pot = [l for l in lines if l.startswith("potential:")][0]
    assert pot.endswith("NOT FOUND")


def test_empty_values_and_bad_numbers(prm_dir):
    from ctbb_icd_amd import configs
    txt = configs.prm_text("tiny")
    txt = txt.replace("nx: 24", "nx:        # empty").replace("Nrows: 8", "Nrows: 8.5")
    txt = txt.replace("penalty: quadratic", "penalty: huber")
    p = prm_dir / "bad.yaml"
    p.write_text(txt)
    out = run_configure(p)
    assert [l for l in out.splitlines() if l.startswith("nx:")][0].endswith("NOT FOUND")
    assert [l for l in out.splitlines() if l.startswith("Nrows:")][0].endswith("NOT FOUND")
    assert "nx was not explicitly set in parameter file. Using default value: 512" in out
    assert 'Requesting penalty function, "huber", is invalid.' in out
    assert "RC -1" in out


def test_missing_file(prm_dir):
    out = run_configure(prm_dir / "does_not_exist.yaml")
    assert "Parameter file not found. Exiting" in out
    assert "RC -4" in out
