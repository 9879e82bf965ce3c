"""The C-ABI library loads (no GPU needed) and exports every declared entry."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(ctbb_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol(ct):
    import ctypes
    L = ct.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = []
    for n in names:
        try:
            getattr(L, n)
        except AttributeError:
            missing.append(n)
    assert not missing, missing


def test_version_and_last_error(ct):
    assert b"gfx950" in ct.lib().ctbb_version()
    # a host-side error path: a missing prm file
    rc = ct.lib().ctbb_configure(b"/nonexistent/prm.yaml", None, None, None, None, None, None)
    assert rc != 0
    assert ct.lib().ctbb_last_error()


def test_device_entry_points_reject_null(ct):
    import ctypes
    assert ct.lib().ctbb_ssm_build(None, None) == -1
    assert ct.lib().ctbb_recon_iterate(None, 1, None) == -1


def test_empty_slab_is_rejected(ct, tmp_path):
    """z_start == z_end gives no slices; the matrix build refuses the
    geometry before touching the device."""
    from ctbb_icd_amd import configs
    s = ct.Scan(configs.write_prm("tiny", tmp_path / "empty.yaml", z_start=-1.5, z_end=-1.5))
    assert s.geom.nz == 0
    with pytest.raises(ct.CtbbError, match="positive"):
        ct.SSM.build(s.geom)


def test_buffer_sizes_are_checked_before_the_abi(ct, make_scan):
    """The binding refuses buffers whose size disagrees with the geometry
    before any pointer reaches the library (which trusts the sizes)."""
    import ctypes
    import numpy as np
    s = make_scan("tiny")
    g = s.geom
    m = ct.SSM(ctypes.c_void_p(), g)          # never reaches the library
    x = np.zeros(g.nvox)
    raw = np.zeros(g.data_size, np.float32)
    with pytest.raises(ValueError, match="volume"):
        m.forward_project(s.slice_index, x[:-1])
    with pytest.raises(ValueError, match="slice_index"):
        m.forward_project(s.slice_index[:-1], x)
    with pytest.raises(ValueError, match="raw"):
        m.fbp(raw[:-1], s.slice_index)
    with pytest.raises(ValueError, match="raw"):
        ct.Recon(m, raw[1:], None, s.slice_index, s.opts())
    with pytest.raises(ValueError, match="weights"):
        ct.Recon(m, raw, raw[1:], s.slice_index, s.opts())
    with pytest.raises(ValueError, match="volume"):
        ct.cost(m, raw, None, s.slice_index, s.opts(), x[1:])
    with pytest.raises(ValueError, match="tube_angles"):
        ct.rotate_slices(g, s.tube_angles[1:], s.slice_index, x.astype(np.float32), 1)
    with pytest.raises(ValueError, match="alpha"):
        ct.penalty_steps(g, s.opts(), x, np.arange(3), np.zeros(2), np.zeros(3))
    with pytest.raises(ValueError, match="idx"):
        ct.SSM.upload(g, np.zeros(g.nx * g.ny + 1, np.int64), np.zeros(1, np.int32),
                      np.zeros(0, np.float32))
