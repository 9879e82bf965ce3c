"""FileType-0 binary raw frames (load_raw, src/setup.cpp:271-302, 335-344).

The frame reader of the reference lives in the absent ctbb_read library;
the file layout tested here (RawOffset header bytes, then Readings frames of
n_channels x Nrows float32 samples, rows fastest) is the plain reading of its
call sites, where the reference's two frame transposes cancel.
"""
import numpy as np
import pytest


def _write(path, raw, offset):
    with open(path, "wb") as f:
        f.write(b"\x5a" * offset)
        f.write(np.ascontiguousarray(raw, np.float32).tobytes())


@pytest.mark.parametrize("offset", [0, 1024])
def test_binary_frames_round_trip(make_scan, ct, tmp_path, offset):
    s = make_scan("tiny_ffs")
    raw, _ = s.synthesize(use_gpu=False)
    f = tmp_path / f"frames_{offset}.bin"
    _write(f, raw, offset)
    from ctbb_icd_amd import configs
    s2 = ct.Scan(configs.write_prm("tiny_ffs", tmp_path / f"bin_{offset}.yaml", synthetic=0,
                                   sinogram_path=str(f), RawOffset=offset))
    got = s2.read_raw()
    assert np.array_equal(got, raw)
    # sample (view, channel, row) sits at row + Nrows * (channel + n_channels * view)
    g = s2.geom
    v, c, r = 5, 7, 3
    assert got[r + g.n_rows * (c + g.n_channels * v)] == raw.reshape(
        g.readings, g.n_channels, g.n_rows)[v, c, r]


def test_short_missing_and_unsupported_files(make_scan, ct, tmp_path):
    from ctbb_icd_amd import configs
    s = make_scan("tiny")
    raw, _ = s.synthesize(use_gpu=False)
    f = tmp_path / "short.bin"
    _write(f, raw[: raw.size // 2], 0)
    for over, code in [(dict(sinogram_path=str(f)), "fewer than"),
                       (dict(sinogram_path=str(tmp_path / "none.bin")), "not found"),
                       # PTR angle/table conventions are refused when the
                       # scan is configured, before any frame is read
                       (dict(sinogram_path=str(f), FileType=1), "slice table rejected")]:
        p = configs.write_prm("tiny", tmp_path / f"bad_{len(code)}.yaml", synthetic=0, **over)
        with pytest.raises(Exception, match=code):
            ct.Scan(p).read_raw()
