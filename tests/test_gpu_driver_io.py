"""matrix.bin interop (src/generate_system_matrix.cpp:212-303, read at
src/icd_iteration.cpp:164-179) and the driver end to end (src/main.cpp:89-128,
output layout of src/setup.cpp:422-438)."""
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(ct):
    ct.init(0)
    return ct


def _parse_matrix_bin(path, ncol):
    """Independent reader of the reference layout: per column (ix fastest)
    uint64 nnz, then nnz x {int32 index, float32 value}; no header."""
    buf = open(path, "rb").read()
    rec = np.dtype([("q", "<i4"), ("v", "<f4")])
    cp = np.zeros(ncol + 1, np.int64)
    qs, vs, o = [], [], 0
    for c in range(ncol):
        n = int(np.frombuffer(buf, "<u8", 1, o)[0])
        o += 8
        r = np.frombuffer(buf, rec, n, o)
        o += 8 * n
        qs.append(r["q"])
        vs.append(r["v"])
        cp[c + 1] = cp[c] + n
    assert o == len(buf)
    return cp, np.concatenate(qs), np.concatenate(vs)


def test_matrix_bin_layout_and_round_trip(gpu, make_scan, orc, tmp_path):
    s = make_scan("tiny_ffs")
    g = s.geom
    m = gpu.SSM.build(g)
    path = tmp_path / "matrix.bin"
    m.write_file(str(path))
    cp, q, v = _parse_matrix_bin(path, g.nx * g.ny)
    ocp, oq, ov = orc.ssm_build(g)
    assert np.array_equal(cp, ocp) and np.array_equal(q, oq) and np.array_equal(v, ov)
    m2 = gpu.SSM.read_file(g, str(path))
    a, b = m.download(), m2.download()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_driver_end_to_end(gpu, ct, tmp_path):
    from ctbb_icd_amd import configs
    out = tmp_path / "out"
    (out / "reconstructions").mkdir(parents=True)
    prm = configs.write_prm("tiny_ffs", tmp_path / "scan.yaml", output_dir=str(out),
                            num_iterations=2)
    driver = ROOT / "ctbb_icd_amd" / "ctbb_icd"
    runs = []
    for attempt in range(2):   # second run reuses output_dir/matrix.bin
        p = subprocess.run([str(driver), str(prm)], capture_output=True, text=True,
                           timeout=120)
        print(p.stdout[-600:], flush=True)
        assert p.returncode == 0, p.stdout + p.stderr
        runs.append(np.fromfile(out / "reconstruction.img", np.float32))
    assert "Existing matrix file FOUND" in p.stdout
    assert (out / "matrix.bin").exists()
    s = ct.Scan(prm)
    g = s.geom
    assert runs[0].size == g.nvox
    assert np.array_equal(runs[0], runs[1])
    # the same steps through the library
    raw, w = s.synthesize(use_gpu=True)
    m = ct.SSM.build(g)
    r = ct.Recon(m, raw, w, s.slice_index, s.opts())
    r.set_volume(np.zeros(g.nvox, np.float32))
    r.iterate(1)
    r.iterate(1)
    x = r.volume()
    dumped = np.fromfile(out / "reconstructions" / "iteration2.rcn", np.float64)
    assert np.array_equal(dumped, x.astype(np.float64))
    assert np.fromfile(out / "reconstructions" / "iteration0.rcn", np.float64).size == g.nvox
    # zero start: the FP64 sinogram estimate A x0 is zero (src/icd_iteration.cpp:127-133)
    sino = np.fromfile(out / "reconstructions" / "sino_estimation.rcn", np.float64)
    assert sino.size == g.data_size and not sino.any()
    fixed = ct.rotate_slices(g, s.tube_angles, s.slice_index, x, 1)
    assert np.array_equal(runs[1], fixed)


def test_driver_reads_binary_frames(gpu, ct, tmp_path):
    """The driver on a FileType-0 binary file (synthetic: 0) reconstructs the
    same volume as the library does from the same frames (unweighted, as the
    reference is for binary files)."""
    from ctbb_icd_amd import configs
    out = tmp_path / "outb"
    (out / "reconstructions").mkdir(parents=True)
    s0 = ct.Scan(configs.write_prm("tiny_ffs", tmp_path / "gen.yaml", synthetic_weights=0))
    raw, _ = s0.synthesize(use_gpu=True)
    frames = tmp_path / "frames.bin"
    with open(frames, "wb") as f:
        f.write(bytes(256))
        f.write(raw.astype(np.float32).tobytes())
    prm = configs.write_prm("tiny_ffs", tmp_path / "bin.yaml", output_dir=str(out),
                            num_iterations=2, synthetic=0, sinogram_path=str(frames),
                            RawOffset=256)
    driver = ROOT / "ctbb_icd_amd" / "ctbb_icd"
    p = subprocess.run([str(driver), str(prm)], capture_output=True, text=True, timeout=120)
    print(p.stdout[-400:], flush=True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "binary frames" in p.stdout
    vol = np.fromfile(out / "reconstruction.img", np.float32)
    s = ct.Scan(prm)
    g = s.geom
    m = ct.SSM.build(g)
    r = ct.Recon(m, s.read_raw(), None, s.slice_index, s.opts())
    r.set_volume(np.zeros(g.nvox, np.float32))
    r.iterate(1)
    r.iterate(1)
    fixed = ct.rotate_slices(g, s.tube_angles, s.slice_index, r.volume(), 1)
    assert np.array_equal(vol, fixed)


def test_driver_fbp_start(gpu, ct, tmp_path):
    """wfbp_initialize: 1 starts from the GPU FBP image (the external wFBP
    program of the reference is replaced); the run completes and differs
    from the zero start."""
    from ctbb_icd_amd import configs
    outs = []
    for wfbp in (0, 1):
        out = tmp_path / f"out_fbp{wfbp}"
        (out / "reconstructions").mkdir(parents=True)
        prm = configs.write_prm("tiny_ffs", tmp_path / f"fbp{wfbp}.yaml", output_dir=str(out),
                                num_iterations=1, wfbp_initialize=wfbp)
        p = subprocess.run([str(ROOT / "ctbb_icd_amd" / "ctbb_icd"), str(prm)],
                           capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stdout + p.stderr
        if wfbp:
            assert "Initialized from the GPU FBP image" in p.stdout
        outs.append(np.fromfile(out / "reconstruction.img", np.float32))
    assert not np.array_equal(outs[0], outs[1])


def test_driver_initial_volume_from_file(gpu, ct, orc, tmp_path):
    """initial_recon_path (src/initialize_recon_volume.cpp:51-76): a float32
    volume in the rotated frame starts the run (its projection is formed
    first); a missing or truncated file stops the driver with the
    reference's messages."""
    from ctbb_icd_amd import configs
    out = tmp_path / "outf"
    (out / "reconstructions").mkdir(parents=True)
    s0 = ct.Scan(configs.write_prm("tiny_ffs", tmp_path / "s0.yaml"))
    g = s0.geom
    x0 = (s0.phantom() * 0.9).astype(np.float32)
    start = tmp_path / "start.img"
    x0.tofile(start)
    driver = ROOT / "ctbb_icd_amd" / "ctbb_icd"
    prm = configs.write_prm("tiny_ffs", tmp_path / "f.yaml", output_dir=str(out),
                            num_iterations=1, initial_recon_path=str(start))
    p = subprocess.run([str(driver), str(prm)], capture_output=True, text=True, timeout=120)
    print(p.stdout[-400:], flush=True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "Loading initial reconstruction from file" in p.stdout
    it0 = np.fromfile(out / "reconstructions" / "iteration0.rcn", np.float64)
    assert np.array_equal(it0, x0.astype(np.float64))
    # the sinogram estimate is the FP64 projection of the start volume
    sino = np.fromfile(out / "reconstructions" / "sino_estimation.rcn", np.float64)
    y = orc.forward_project(g, orc.ssm_build(g), s0.slice_index, x0.astype(np.float64))
    assert np.abs(sino - y).max() <= 1e-12 * np.abs(y).max()
    # the same run through the library
    s = ct.Scan(prm)
    raw, w = s.synthesize(use_gpu=True)
    m = ct.SSM.build(g)
    r = ct.Recon(m, raw, w, s.slice_index, s.opts())
    r.set_volume(x0)
    r.iterate(1)
    fixed = ct.rotate_slices(g, s.tube_angles, s.slice_index, r.volume(), 1)
    assert np.array_equal(np.fromfile(out / "reconstruction.img", np.float32), fixed)
    # truncated file
    short = tmp_path / "short.img"
    x0[: g.nvox // 2].tofile(short)
    prm2 = configs.write_prm("tiny_ffs", tmp_path / "f2.yaml", output_dir=str(out),
                             num_iterations=1, initial_recon_path=str(short))
    p = subprocess.run([str(driver), str(prm2)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 1
    assert "File was not read properly." in p.stdout
    assert "Could not initialize reconstruction volume from given file. Exiting" in p.stdout
    # missing file
    prm3 = configs.write_prm("tiny_ffs", tmp_path / "f3.yaml", output_dir=str(out),
                             num_iterations=1, initial_recon_path=str(tmp_path / "none.img"))
    p = subprocess.run([str(driver), str(prm3)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 1
    assert "Specified initial reconstruction file not found. Exiting" in p.stdout
