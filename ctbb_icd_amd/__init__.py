"""ctbb_icd_amd -- MI355X (gfx950) reconstruction hot path of CTBB_ICD.

Python access to libctbb_icd.so through its C ABI (include/ctbb_icd.h,
include/ctbb_host.h).  The library is the product; this module only binds it.
Importing works without a GPU (host-side entry points run on the CPU); every
device entry point fails loudly when no HIP device is present.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("CTBB_ICD_LIBRARY", str(PKG / "libctbb_icd.so")))

CTBB_SSM_FP32 = 0
CTBB_SSM_FP16 = 1
PENALTY_QUADRATIC = 0
PENALTY_EDGE_PRESERVING = 1

_i64, _f64, _i32 = C.c_int64, C.c_double, C.c_int32


class Geom(C.Structure):
    _fields_ = [
        ("nx", _i64), ("ny", _i64), ("nz", _i64), ("n_channels", _i64),
        ("n_rows", _i64), ("readings", _i64), ("n_views_sm", _i64),
        ("zffs", _i64), ("phiffs", _i64), ("table_direction", _i64),
        ("tube_angle_increment", _f64), ("tube_z_increment", _f64),
        ("focal_spot_radius", _f64), ("source_detector_distance", _f64),
        ("center_channel_non_ffs", _f64), ("center_row", _f64),
        ("transaxial_detector_spacing", _f64), ("axial_detector_spacing", _f64),
        ("fan_angle_increment", _f64), ("anode_angle", _f64),
        ("coll_slicewidth", _f64), ("center_voxel_x", _f64),
        ("center_voxel_y", _f64), ("center_voxel_z", _f64),
        ("voxel_size_x", _f64), ("voxel_size_y", _f64), ("voxel_size_z", _f64),
        ("acquisition_fov", _f64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}

    @property
    def data_size(self) -> int:
        return self.readings * self.n_channels * self.n_rows

    @property
    def nvox(self) -> int:
        return self.nx * self.ny * self.nz


class RunParams(C.Structure):
    _fields_ = [
        ("num_iterations", _i64), ("penalty", _i32), ("wfbp_initialize", _i32),
        ("lambda_", _f64), ("delta", _f64), ("views_per_slice", _i64),
        ("num_views_per_turn", _i64), ("z_first_view", _f64), ("z_start", _f64),
        ("z_end", _f64), ("beam_width", _f64), ("beam_width_at_roi_edge", _f64),
        ("recon_fov", _f64), ("synthetic", _i32), ("use_weights", _i32),
        ("ssm_fp16", _i32), ("neighbours", _i32), ("resync_every", _i32),
        ("state_fp64", _i32), ("group_spacing", _i64),
    ]


class IcdOpts(C.Structure):
    _fields_ = [
        ("penalty", _i32), ("neighbours", _i32), ("lambda_", _f64),
        ("delta", _f64), ("group_spacing", _i64), ("resync_every", _i32),
        ("tile", _i32), ("plain_curvature", _i32), ("state_fp64", _i32),
    ]


class IcdStats(C.Structure):
    _fields_ = [
        ("iterations", _i64), ("voxel_updates", _i64), ("entry_visits", _i64),
        ("sweep_ms", _f64), ("resync_ms", _f64), ("cost", _f64),
        ("sweep_launches", _i64), ("ssm_entries", _i64),
        ("gather_ms_avg", _f64), ("scatter_ms_avg", _f64), ("solve_ms_avg", _f64),
        ("tile", _i64), ("group_spacing", _i64),
    ]


class CtbbError(RuntimeError):
    pass


_lib = None


def lib() -> C.CDLL:
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise CtbbError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                        "or python ctbb_icd_amd/build.py")
    L = C.CDLL(str(LIB_PATH))
    P = C.c_void_p
    sigs = {
        "ctbb_last_error": ([], C.c_char_p),
        "ctbb_version": ([], C.c_char_p),
        "ctbb_init": ([C.c_int], C.c_int),
        "ctbb_set_verbose": ([C.c_int], None),
        "ctbb_copy_probe": ([_i64, C.POINTER(_f64)], C.c_int),
        "ctbb_configure": ([C.c_char_p, C.POINTER(Geom), C.POINTER(RunParams), P, P, P, P], C.c_int),
        "ctbb_synthesize": ([C.c_char_p, P, P, C.c_int], C.c_int),
        "ctbb_read_raw": ([C.c_char_p, P], C.c_int),
        "ctbb_phantom_volume": ([C.c_char_p, P], C.c_int),
        "ctbb_ssm_build": ([C.POINTER(Geom), C.POINTER(P)], C.c_int),
        "ctbb_ssm_to_fp16": ([P], C.c_int),
        "ctbb_ssm_free": ([P], C.c_int),
        "ctbb_ssm_info": ([P, P, P, P], C.c_int),
        "ctbb_ssm_download": ([P, P, P, P], C.c_int),
        "ctbb_ssm_download_cols": ([P, _i64, _i64, P, P, P], C.c_int),
        "ctbb_ssm_upload": ([C.POINTER(Geom), P, P, P, C.POINTER(P)], C.c_int),
        "ctbb_ssm_write_file": ([P, C.c_char_p], C.c_int),
        "ctbb_ssm_read_file": ([C.POINTER(Geom), C.c_char_p, C.POINTER(P)], C.c_int),
        "ctbb_ssm_group_spacing": ([P, P, _i64, _f64, P], C.c_int),
        "ctbb_forward_project": ([P, P, P, P], C.c_int),
        "ctbb_fbp": ([P, P, P, P], C.c_int),
        "ctbb_recon_create": ([P, P, P, P, C.POINTER(IcdOpts), C.POINTER(P)], C.c_int),
        "ctbb_recon_free": ([P], C.c_int),
        "ctbb_recon_set_volume": ([P, P], C.c_int),
        "ctbb_recon_get_volume": ([P, P], C.c_int),
        "ctbb_recon_set_volume64": ([P, P], C.c_int),
        "ctbb_recon_get_volume64": ([P, P], C.c_int),
        "ctbb_recon_iterate": ([P, C.c_int, C.POINTER(IcdStats)], C.c_int),
        "ctbb_recon_resync": ([P], C.c_int),
        "ctbb_recon_cost": ([P, P], C.c_int),
        "ctbb_recon_cost_running": ([P, P], C.c_int),
        "ctbb_recon_get_error": ([P, P], C.c_int),
        "ctbb_cost": ([P, P, P, P, C.POINTER(IcdOpts), P, P], C.c_int),
        "ctbb_cost_noalloc": ([P, P, P, P, C.POINTER(IcdOpts), P, P, P], C.c_int),
        "ctbb_rotate_slices": ([C.POINTER(Geom), P, P, P, C.c_int], C.c_int),
        "ctbb_penalty_steps": ([C.POINTER(Geom), C.POINTER(IcdOpts), P, C.c_int, P, _i64, P, P,
                                P, P], C.c_int),
        "ctbb_unet_create": ([C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64,
                              C.POINTER(P)], C.c_int),
        "ctbb_unet_free": ([P], C.c_int),
        "ctbb_unet_num_params": ([P, C.POINTER(_i64)], C.c_int),
        "ctbb_unet_get_params": ([P, P], C.c_int),
        "ctbb_unet_set_params": ([P, P], C.c_int),
        "ctbb_unet_grad": ([P, P, P, C.c_int, C.POINTER(_f64), P], C.c_int),
        "ctbb_unet_adam_step": ([P, _f64, _f64, _f64, _f64], C.c_int),
        "ctbb_unet_infer": ([P, P, C.c_int, P], C.c_int),
        "ctbb_unet_last_timing": ([P, C.POINTER(_f64), C.POINTER(_f64)], C.c_int),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def exported_symbols() -> list[str]:
    return [n for n in dir(lib()) if n.startswith("ctbb_")]


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().ctbb_last_error().decode(errors="replace")
        raise CtbbError(f"ctbb error {rc}: {msg}")


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _need(a: np.ndarray | None, n: int, what: str) -> None:
    """Size check before a buffer crosses the C ABI: the library reads and
    writes as many elements as the geometry implies."""
    if a is not None and a.size != n:
        raise ValueError(f"{what}: {a.size} elements, expected {n}")


def set_verbose(on: bool) -> None:
    lib().ctbb_set_verbose(1 if on else 0)


# ---------------------------------------------------------------------------
# Scan set-up (host side)

class Scan:
    """A configured scan: geometry, run parameters and slice table of a prm."""

    def __init__(self, prm_path: str | Path):
        self.prm_path = str(prm_path)
        g, rp = Geom(), RunParams()
        check(lib().ctbb_configure(self.prm_path.encode(), C.byref(g), C.byref(rp),
                                   None, None, None, None))
        self.tube_angles = np.zeros(g.readings, np.float64)
        self.table_positions = np.zeros(g.readings, np.float64)
        self.slice_locations = np.zeros(g.nz, np.float64)
        self.slice_index = np.zeros(g.nz, np.int64)
        check(lib().ctbb_configure(self.prm_path.encode(), C.byref(g), C.byref(rp),
                                   _p(self.tube_angles), _p(self.table_positions),
                                   _p(self.slice_locations), _p(self.slice_index)))
        self.geom, self.run = g, rp
        self.tile = 0
        self.plain_curvature = 0
        for line in open(self.prm_path):
            key = line.split(":")[0].strip()
            if key in ("tile", "plain_curvature"):
                setattr(self, key, int(line.split(":", 1)[1].split("#")[0]))

    def synthesize(self, use_gpu: bool = False, weights: bool | None = None):
        g = self.geom
        raw = np.zeros(g.data_size, np.float32)
        want_w = self.run.use_weights if weights is None else weights
        w = np.zeros(g.data_size, np.float32) if want_w else None
        check(lib().ctbb_synthesize(self.prm_path.encode(), _p(raw), _p(w),
                                    1 if use_gpu else 0))
        return raw, w

    def read_raw(self) -> np.ndarray:
        """FileType-0 binary frames of the prm's sinogram_path."""
        raw = np.zeros(self.geom.data_size, np.float32)
        check(lib().ctbb_read_raw(self.prm_path.encode(), _p(raw)))
        return raw

    def phantom(self) -> np.ndarray:
        v = np.zeros(self.geom.nvox, np.float32)
        check(lib().ctbb_phantom_volume(self.prm_path.encode(), _p(v)))
        return v

    def opts(self, **kw) -> IcdOpts:
        o = IcdOpts()
        o.penalty = kw.get("penalty", self.run.penalty)
        o.neighbours = kw.get("neighbours", self.run.neighbours or 8)
        o.lambda_ = kw.get("lam", self.run.lambda_)
        o.delta = kw.get("delta", self.run.delta)
        o.group_spacing = kw.get("group_spacing", self.run.group_spacing)
        o.resync_every = kw.get("resync_every", self.run.resync_every)
        o.tile = kw.get("tile", self.tile)
        o.plain_curvature = kw.get("plain_curvature", self.plain_curvature)
        o.state_fp64 = kw.get("state_fp64", self.run.state_fp64)
        return o


# ---------------------------------------------------------------------------
# Device objects

class SSM:
    """Device-resident stored system matrix."""

    def __init__(self, handle: C.c_void_p, geom: Geom):
        self.h = handle
        self.geom = geom

    @classmethod
    def build(cls, geom: Geom) -> "SSM":
        h = C.c_void_p()
        check(lib().ctbb_ssm_build(C.byref(geom), C.byref(h)))
        return cls(h, geom)

    @classmethod
    def upload(cls, geom: Geom, col_ptr, idx, val) -> "SSM":
        h = C.c_void_p()
        col_ptr = np.ascontiguousarray(col_ptr, np.int64)
        idx = np.ascontiguousarray(idx, np.int32)
        val = np.ascontiguousarray(val, np.float32)
        _need(col_ptr, geom.nx * geom.ny + 1, "col_ptr")
        _need(idx, int(col_ptr[-1]), "idx")
        _need(val, int(col_ptr[-1]), "val")
        check(lib().ctbb_ssm_upload(C.byref(geom), _p(col_ptr), _p(idx), _p(val),
                                    C.byref(h)))
        return cls(h, geom)

    @classmethod
    def read_file(cls, geom: Geom, path: str) -> "SSM":
        h = C.c_void_p()
        check(lib().ctbb_ssm_read_file(C.byref(geom), str(path).encode(), C.byref(h)))
        return cls(h, geom)

    def info(self):
        nnz, prec, ms = C.c_int64(), C.c_int(), C.c_double()
        check(lib().ctbb_ssm_info(self.h, C.byref(nnz), C.byref(prec), C.byref(ms)))
        return nnz.value, prec.value, ms.value

    @property
    def nnz(self) -> int:
        return self.info()[0]

    def to_fp16(self) -> None:
        check(lib().ctbb_ssm_to_fp16(self.h))

    def download(self):
        ncol = self.geom.nx * self.geom.ny
        nnz = self.nnz
        cp = np.zeros(ncol + 1, np.int64)
        idx = np.zeros(max(nnz, 1), np.int32)
        val = np.zeros(max(nnz, 1), np.float32)
        check(lib().ctbb_ssm_download(self.h, _p(cp), _p(idx), _p(val)))
        return cp, idx[:nnz], val[:nnz]

    def download_cols(self, c0: int, c1: int):
        cp = np.zeros(c1 - c0 + 1, np.int64)
        check(lib().ctbb_ssm_download_cols(self.h, c0, c1, _p(cp), None, None))
        n = int(cp[-1])
        idx = np.zeros(max(n, 1), np.int32)
        val = np.zeros(max(n, 1), np.float32)
        check(lib().ctbb_ssm_download_cols(self.h, c0, c1, _p(cp), _p(idx), _p(val)))
        return cp, idx[:n], val[:n]

    def write_file(self, path: str) -> None:
        check(lib().ctbb_ssm_write_file(self.h, str(path).encode()))

    def group_spacing(self, slice_index: np.ndarray, max_spacing: int = 0,
                      max_conflict_fraction: float = 0.0) -> int:
        s = C.c_int64()
        si = np.ascontiguousarray(slice_index, np.int64)
        _need(si, self.geom.nz, "slice_index")
        check(lib().ctbb_ssm_group_spacing(self.h, _p(si), max_spacing,
                                           max_conflict_fraction, C.byref(s)))
        return s.value

    def forward_project(self, slice_index: np.ndarray, x: np.ndarray) -> np.ndarray:
        si = np.ascontiguousarray(slice_index, np.int64)
        x = np.ascontiguousarray(x, np.float64)
        _need(si, self.geom.nz, "slice_index")
        _need(x, self.geom.nvox, "volume")
        y = np.zeros(self.geom.data_size, np.float64)
        check(lib().ctbb_forward_project(self.h, _p(si), _p(x), _p(y)))
        return y

    def fbp(self, raw: np.ndarray, slice_index: np.ndarray) -> np.ndarray:
        """Filtered backprojection start image (rotating frame, FP32)."""
        si = np.ascontiguousarray(slice_index, np.int64)
        raw = np.ascontiguousarray(raw, np.float32)
        g = self.geom
        _need(raw, g.data_size, "raw")
        _need(si, g.nz, "slice_index")
        x = np.zeros(g.nvox, np.float32)
        check(lib().ctbb_fbp(self.h, _p(raw), _p(si), _p(x)))
        return x

    def free(self) -> None:
        if self.h:
            lib().ctbb_ssm_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Recon:
    """Device-resident ICD reconstruction context."""

    def __init__(self, ssm: SSM, raw: np.ndarray, weights, slice_index, opts: IcdOpts):
        self.ssm = ssm
        self.geom = ssm.geom
        self._raw = np.ascontiguousarray(raw, np.float32)
        self._w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        self._si = np.ascontiguousarray(slice_index, np.int64)
        _need(self._raw, self.geom.data_size, "raw")
        _need(self._w, self.geom.data_size, "weights")
        _need(self._si, self.geom.nz, "slice_index")
        self.opts = opts
        self.h = None
        h = C.c_void_p()
        check(lib().ctbb_recon_create(ssm.h, _p(self._raw), _p(self._w), _p(self._si),
                                      C.byref(opts), C.byref(h)))
        self.h = h

    def set_volume(self, vol: np.ndarray) -> None:
        vol = np.ascontiguousarray(vol, np.float32)
        _need(vol, self.geom.nvox, "volume")
        check(lib().ctbb_recon_set_volume(self.h, _p(vol)))

    def volume(self) -> np.ndarray:
        v = np.zeros(self.geom.nvox, np.float32)
        check(lib().ctbb_recon_get_volume(self.h, _p(v)))
        return v

    def volume64(self) -> np.ndarray:
        v = np.zeros(self.geom.nvox, np.float64)
        check(lib().ctbb_recon_get_volume64(self.h, _p(v)))
        return v

    def set_volume64(self, vol: np.ndarray) -> None:
        vol = np.ascontiguousarray(vol, np.float64)
        _need(vol, self.geom.nvox, "volume")
        check(lib().ctbb_recon_set_volume64(self.h, _p(vol)))

    def error(self) -> np.ndarray:
        e = np.zeros(self.geom.data_size, np.float32)
        check(lib().ctbb_recon_get_error(self.h, _p(e)))
        return e

    def iterate(self, n: int) -> IcdStats:
        st = IcdStats()
        check(lib().ctbb_recon_iterate(self.h, n, C.byref(st)))
        return st

    def resync(self) -> None:
        check(lib().ctbb_recon_resync(self.h))

    def cost(self) -> float:
        c = C.c_double()
        check(lib().ctbb_recon_cost(self.h, C.byref(c)))
        return c.value

    def cost_running(self) -> float:
        """Cost from the kept FP32 error sinogram (no projection)."""
        c = C.c_double()
        check(lib().ctbb_recon_cost_running(self.h, C.byref(c)))
        return c.value

    def free(self) -> None:
        if self.h:
            lib().ctbb_recon_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def cost(ssm: "SSM", raw, weights, slice_index, opts: IcdOpts, volume) -> float:
    """FP64 cost of a volume (ctbb_cost), no reconstruction context needed."""
    raw = np.ascontiguousarray(raw, np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    si = np.ascontiguousarray(slice_index, np.int64)
    v = np.ascontiguousarray(volume, np.float32)
    g = ssm.geom
    _need(raw, g.data_size, "raw")
    _need(w, g.data_size, "weights")
    _need(si, g.nz, "slice_index")
    _need(v, g.nvox, "volume")
    c = C.c_double()
    check(lib().ctbb_cost(ssm.h, _p(raw), _p(w), _p(si), C.byref(opts), _p(v), C.byref(c)))
    return c.value


def rotate_slices(geom: Geom, tube_angles, slice_index, volume, direction: int) -> np.ndarray:
    v = np.ascontiguousarray(volume, np.float32).copy()
    ta = np.ascontiguousarray(tube_angles, np.float64)
    si = np.ascontiguousarray(slice_index, np.int64)
    _need(v, geom.nvox, "volume")
    _need(ta, geom.readings, "tube_angles")
    _need(si, geom.nz, "slice_index")
    check(lib().ctbb_rotate_slices(C.byref(geom), _p(ta), _p(si), _p(v), direction))
    return v


def penalty_steps(geom: Geom, opts: IcdOpts, x, q, alpha, beta,
                  volume_fp32: bool = False, slice_index=None) -> np.ndarray:
    """Device penalty steps (ctbb_penalty_steps); volume_fp32: evaluate on x
    rounded to FP32, as the sweep holds the volume; slice_index: needed for
    the 26-neighbour penalty."""
    si = None if slice_index is None else np.ascontiguousarray(slice_index, np.int64)
    _need(si, geom.nz, "slice_index")
    x = np.ascontiguousarray(x, np.float64)
    q = np.ascontiguousarray(q, np.int64)
    a = np.ascontiguousarray(alpha, np.float64)
    b = np.ascontiguousarray(beta, np.float64)
    _need(x, geom.nvox, "volume")
    _need(a, q.size, "alpha")
    _need(b, q.size, "beta")
    du = np.zeros(len(q), np.float64)
    check(lib().ctbb_penalty_steps(C.byref(geom), C.byref(opts), _p(si), int(volume_fp32),
                                   _p(x), len(q), _p(q),
                                   _p(a), _p(b), _p(du)))
    return du


class UNet:
    """The warm-start network (include/ctbb_unet.h): BF16 MFMA convolutions,
    FP32 master weights and Adam, all on the device."""

    def __init__(self, levels: int = 4, base: int = 32, max_batch: int = 8,
                 height: int = 256, width: int = 256, seed: int = 1):
        self.levels, self.base, self.max_batch = levels, base, max_batch
        self.height, self.width = height, width
        h = C.c_void_p()
        check(lib().ctbb_unet_create(levels, base, max_batch, height, width, seed,
                                     C.byref(h)))
        self.h = h
        n = _i64()
        check(lib().ctbb_unet_num_params(self.h, C.byref(n)))
        self.num_params = n.value

    def params(self) -> np.ndarray:
        p = np.zeros(self.num_params, np.float32)
        check(lib().ctbb_unet_get_params(self.h, _p(p)))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params
        check(lib().ctbb_unet_set_params(self.h, _p(p)))

    def grad(self, inp, target, want_grads: bool = False):
        inp = np.ascontiguousarray(inp, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        b = inp.shape[0]
        assert inp.shape == (b, 3, self.height, self.width)
        assert target.shape == (b, self.height, self.width)
        loss = C.c_double()
        g = np.zeros(self.num_params, np.float32) if want_grads else None
        check(lib().ctbb_unet_grad(self.h, _p(inp), _p(target), b, C.byref(loss), _p(g)))
        return (loss.value, g) if want_grads else loss.value

    def adam_step(self, lr: float, beta1: float = 0.9, beta2: float = 0.999,
                  eps: float = 1e-8) -> None:
        check(lib().ctbb_unet_adam_step(self.h, lr, beta1, beta2, eps))

    def infer(self, inp) -> np.ndarray:
        inp = np.ascontiguousarray(inp, np.float32)
        b = inp.shape[0]
        if inp.shape != (b, 3, self.height, self.width) or not 1 <= b <= self.max_batch:
            raise ValueError(f"input {inp.shape}: expected (b <= {self.max_batch}, 3, "
                             f"{self.height}, {self.width})")
        out = np.zeros((b, self.height, self.width), np.float32)
        check(lib().ctbb_unet_infer(self.h, _p(inp), b, _p(out)))
        return out

    def last_timing(self):
        ms, fl = C.c_double(), C.c_double()
        check(lib().ctbb_unet_last_timing(self.h, C.byref(ms), C.byref(fl)))
        return ms.value, fl.value

    def free(self) -> None:
        if getattr(self, "h", None):
            lib().ctbb_unet_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy_probe(nbytes: int) -> float:
    """Run the calibration copy kernel over nbytes; returns milliseconds."""
    ms = C.c_double()
    check(lib().ctbb_copy_probe(nbytes, C.byref(ms)))
    return ms.value


_device = 0


def init(device: int = 0) -> None:
    """Select the HIP device for the calling host thread (HIP's current
    device is per thread: worker threads call init(current_device()))."""
    global _device
    check(lib().ctbb_init(device))
    _device = device


def current_device() -> int:
    """The device the last init() selected."""
    return _device
