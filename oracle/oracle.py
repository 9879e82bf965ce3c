"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it was
built, to the reference's own penalty/rotation code (oracle/_ref).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU baseline legs (the ICD and the matrix build), as the checker
or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libctbb_ref.so"

_P = C.c_void_p
_i64, _f64 = C.c_int64, C.c_double


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def build() -> None:
    r = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stdout + r.stderr)


_lib = None
_ref = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            build()
        L = C.CDLL(str(ORACLE_SO))
        G = _P  # pointer to ctbb_geom
        sig = {
            "orc_ssm_count": ([G, _P], _i64),
            "orc_ssm_fill": ([G, _P, _P, _P], C.c_int),
            "orc_ssm_count_views": ([G, _i64, _i64, _P], _i64),
            "orc_ssm_fill_views": ([G, _i64, _i64, _P, _P, _P], C.c_int),
            "orc_forward_project": ([G, _P, _P, _P, _P, _P, _P], None),
            "orc_quadratic": ([G, _i64, _f64, _f64, _f64, _P], _f64),
            "orc_edge_preserving": ([G, _i64, _f64, _f64, _f64, _f64, _P], _f64),
            "orc_icd": ([G, _P, _P, _P, _P, _P, _P, C.c_int, _f64, _f64, _i64,
                         _i64, C.c_int, _P, _P], _i64),
            "orc_icd_ex": ([G, _P, _P, _P, _P, _P, _P, C.c_int, _f64, _f64, _i64,
                            _i64, C.c_int, C.c_int, _P, _P], _i64),
            "orc_cost": ([G, _P, _P, _P, _P, _P, _P, C.c_int, _f64, _f64, _P], _f64),
            "orc_rotate_slices": ([G, _P, _P, _P, C.c_int], None),
            "orc_column_overlaps": ([G, _P, _P, _i64, _i64], C.c_int),
            "orc_icd_ref_omp": ([G, _P, _P, _P, _P, _P, _P, C.c_int, _f64, _f64,
                                 C.c_int, C.c_int, _i64, _i64, _P, _P], _i64),
            "orc_icd_ref_omp_file": ([G, C.c_char_p, _P, _P, _P, C.c_int, _f64, _f64,
                                      C.c_int, C.c_int, _i64, _i64, _P, _P], _i64),
            "orc_icd_trace": ([G, _P, _P, _P, _P, _P, _P, C.c_int, _f64, _f64, _i64,
                               _i64, C.c_int, C.c_int, _P, _P, _P, _i64, _P, _P], _i64),
            "orc_ssm_cols_count": ([G, _i64, _P, _P], _i64),
            "orc_icd_nb": ([G, _P, _P, _P, _P, _P, _P, C.c_int, _f64, _f64, C.c_int, _i64,
                            _i64, C.c_int, C.c_int, _P, _P, _P, _i64, _P, _P], _i64),
            "orc_penalty_step": ([G, _P, C.c_int, C.c_int, _i64, _f64, _f64, _f64, _f64, _P],
                                 _f64),
            "orc_cost_nb": ([G, _P, _P, _P, _P, _P, _P, C.c_int, _f64, _f64, C.c_int, _P],
                            _f64),
            "orc_rot26": ([G, _P, _i64, C.c_int, _i64, _i64, _P, _P], None),
            "orc_ssm_cols_fill": ([G, _i64, _P, _P, _P, _P], None),
        }
        for n, (a, r) in sig.items():
            f = getattr(L, n)
            f.argtypes, f.restype = a, r
        _lib = L
    return _lib


def ref() -> C.CDLL | None:
    """The reference's own penalties/rotation (None if not built here)."""
    global _ref
    if _ref is None and REF_SO.exists():
        L = C.CDLL(str(REF_SO))
        L.ref_quadratic.argtypes = [_i64, _i64, _i64, C.c_int, _f64, _f64, _f64, _P]
        L.ref_quadratic.restype = _f64
        L.ref_edge_preserving.argtypes = [_i64, _i64, _i64, C.c_int, _f64, _f64, _f64,
                                          _f64, _P]
        L.ref_edge_preserving.restype = _f64
        L.ref_rotate_slices.argtypes = [_i64, _i64, _i64, _P, _P, _P, C.c_int]
        L.ref_rotate_slices.restype = None
        _ref = L
    return _ref


def gp(geom):
    return C.cast(C.pointer(geom), C.c_void_p)


def ssm_build(geom):
    ncol = geom.nx * geom.ny
    cnt = np.zeros(ncol, np.int64)
    nnz = lib().orc_ssm_count(gp(geom), _p(cnt))
    if nnz < 0:
        raise RuntimeError("oracle SSM: unsupported geometry")
    cp = np.zeros(ncol + 1, np.int64)
    cp[1:] = np.cumsum(cnt)
    idx = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), np.float32)
    lib().orc_ssm_fill(gp(geom), _p(cp), _p(idx), _p(val))
    return cp, idx[:nnz], val[:nnz]


def ssm_build_views(geom, v0, v1):
    """The ray-driven build over SSM views [v0, v1) only: (ptr, idx, val,
    fill_s), entries numbered as in the full matrix; fill_s is the wall time
    of the pass that traces the rays and writes the entries (the counting
    pass that sizes the arrays is not included)."""
    import time
    ncol = geom.nx * geom.ny
    cnt = np.zeros(ncol, np.int64)
    nnz = lib().orc_ssm_count_views(gp(geom), v0, v1, _p(cnt))
    if nnz < 0:
        raise ValueError(f"oracle SSM: view range [{v0}, {v1}) outside the matrix")
    cp = np.zeros(ncol + 1, np.int64)
    cp[1:] = np.cumsum(cnt)
    idx = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), np.float32)
    t0 = time.perf_counter()
    lib().orc_ssm_fill_views(gp(geom), v0, v1, _p(cp), _p(idx), _p(val))
    return cp, idx[:nnz], val[:nnz], time.perf_counter() - t0


def ssm_cols(geom, cols):
    """Entries of the listed columns (voxel-driven restatement; equal to the
    same columns of ssm_build): (ptr[len(cols)+1], idx, val)."""
    cols = np.ascontiguousarray(cols, np.int64)
    cnt = np.zeros(len(cols), np.int64)
    n = lib().orc_ssm_cols_count(gp(geom), len(cols), _p(cols), _p(cnt))
    ptr = np.zeros(len(cols) + 1, np.int64)
    ptr[1:] = np.cumsum(cnt)
    idx = np.zeros(max(n, 1), np.int32)
    val = np.zeros(max(n, 1), np.float32)
    lib().orc_ssm_cols_fill(gp(geom), len(cols), _p(cols), _p(ptr), _p(idx), _p(val))
    return ptr, idx[:n], val[:n]


def forward_project(geom, ssm, slice_index, x):
    cp, idx, val = ssm
    y = np.zeros(geom.data_size, np.float64)
    lib().orc_forward_project(gp(geom), _p(cp), _p(idx), _p(val),
                              _p(np.ascontiguousarray(slice_index, np.int64)),
                              _p(np.ascontiguousarray(x, np.float64)), _p(y))
    return y


def icd(geom, ssm, raw, weights, slice_index, penalty, lam, delta, spacing,
        n_iter, x=None, yhat=None, tile=8, surrogate=False, neighbours=8):
    """Run n_iter oracle iterations; returns (x, yhat, updates).
    spacing 0: reference order; spacing >= 2: batched z-group order with the
    given tile (tile >= max(nx, ny): single-column batches), optionally with
    the batch surrogate curvature (True / 1: multi-column batches; 2: also
    single-column batches)."""
    cp, idx, val = ssm
    x = np.zeros(geom.nvox, np.float64) if x is None else np.array(x, np.float64)
    if yhat is None:
        yhat = forward_project(geom, ssm, slice_index, x)
    else:
        yhat = np.array(yhat, np.float64)
    raw = np.ascontiguousarray(raw, np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    n = lib().orc_icd_nb(gp(geom), _p(cp), _p(idx), _p(val), _p(raw), _p(w),
                         _p(np.ascontiguousarray(slice_index, np.int64)), penalty, lam,
                         delta, neighbours, spacing, tile, int(surrogate), n_iter, _p(x),
                         _p(yhat), None, 0, None, None)
    return x, yhat, n


def icd_trace(geom, ssm, raw, weights, slice_index, penalty, lam, delta, spacing,
              n_iter, tile, surrogate=1, max_costs=0, x=None, yhat=None):
    """Batched oracle sweep that records the cost after every step.
    Returns (x, yhat, updates, step_costs, shared) with step_costs[0] the
    cost before the first step and shared = samples reached by two voxels of
    one step: (different columns and slices, one column / different slices,
    different columns / one slice)."""
    cp, idx, val = ssm
    x = np.zeros(geom.nvox, np.float64) if x is None else np.array(x, np.float64)
    yhat = forward_project(geom, ssm, slice_index, x) if yhat is None else np.array(yhat)
    raw = np.ascontiguousarray(raw, np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    costs = np.zeros(max(max_costs, 1), np.float64)
    nst = np.zeros(1, np.int64)
    shared = np.zeros(3, np.int64)
    n = lib().orc_icd_trace(gp(geom), _p(cp), _p(idx), _p(val), _p(raw), _p(w),
                            _p(np.ascontiguousarray(slice_index, np.int64)), penalty,
                            lam, delta, spacing, tile, int(surrogate), n_iter, _p(x),
                            _p(yhat), _p(costs) if max_costs else None, max_costs,
                            _p(nst), _p(shared))
    k = min(int(nst[0]) + 1, max_costs)
    return x, yhat, n, costs[:k], tuple(int(v) for v in shared)


def icd_ref_omp(geom, ssm, raw, weights, slice_index, penalty, lam, delta,
                n_iter, threads, row_begin=0, row_end=0, x=None, yhat=None):
    cp, idx, val = ssm
    x = np.zeros(geom.nvox, np.float64) if x is None else np.array(x, np.float64)
    yhat = np.zeros(geom.data_size, np.float64) if yhat is None else np.array(yhat)
    raw = np.ascontiguousarray(raw, np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    n = lib().orc_icd_ref_omp(gp(geom), _p(cp), _p(idx), _p(val), _p(raw), _p(w),
                              _p(np.ascontiguousarray(slice_index, np.int64)),
                              penalty, lam, delta, n_iter, threads, row_begin,
                              row_end, _p(x), _p(yhat))
    return x, yhat, n


def write_band_file(path, geom, ssm, row_begin, row_end):
    """The columns of rows [row_begin, row_end) (x fastest) of a CSC matrix
    (col_ptr, idx, val) in the matrix.bin layout: per column a uint64 nnz and
    nnz {int32 index, float32 value} pairs.  col_ptr may cover the whole
    matrix or be zero outside the band."""
    cp, idx, val = ssm
    pair = np.dtype([("index", "<i4"), ("value", "<f4")])
    with open(path, "wb") as f:
        for c in range(row_begin * geom.nx, row_end * geom.nx):
            a, b = int(cp[c]), int(cp[c + 1])
            f.write(np.uint64(b - a).tobytes())
            rec = np.empty(b - a, pair)
            rec["index"], rec["value"] = idx[a:b], val[a:b]
            f.write(rec.tobytes())


def icd_ref_omp_file(geom, path, raw, weights, slice_index, penalty, lam, delta,
                     n_iter, threads, row_begin, row_end, x=None, yhat=None):
    """orc_icd_ref_omp with the matrix read from `path` (write_band_file of
    the same rows) on every iteration."""
    x = np.zeros(geom.nvox, np.float64) if x is None else np.array(x, np.float64)
    yhat = np.zeros(geom.data_size, np.float64) if yhat is None else np.array(yhat)
    raw = np.ascontiguousarray(raw, np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    n = lib().orc_icd_ref_omp_file(gp(geom), str(path).encode(), _p(raw), _p(w),
                                   _p(np.ascontiguousarray(slice_index, np.int64)),
                                   penalty, lam, delta, n_iter, threads, row_begin,
                                   row_end, _p(x), _p(yhat))
    if n < 0:
        raise IOError(f"matrix file {path} missing or shorter than rows "
                      f"[{row_begin}, {row_end})")
    return x, yhat, n


def cost(geom, ssm, raw, weights, slice_index, penalty, lam, delta, x, neighbours=8):
    cp, idx, val = ssm
    raw = np.ascontiguousarray(raw, np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    return lib().orc_cost_nb(gp(geom), _p(cp), _p(idx), _p(val), _p(raw), _p(w),
                             _p(np.ascontiguousarray(slice_index, np.int64)), penalty,
                             lam, delta, neighbours, _p(np.ascontiguousarray(x, np.float64)))


def penalty_step(geom, slice_index, neighbours, penalty, q, alpha, beta, lam, delta, x):
    return lib().orc_penalty_step(gp(geom), _p(np.ascontiguousarray(slice_index, np.int64)),
                                  neighbours, penalty, q, alpha, beta, lam, delta,
                                  _p(np.ascontiguousarray(x, np.float64)))


def rot26(geom, slice_index, k, i, j, inverse=False):
    io, jo = np.zeros(1, np.int64), np.zeros(1, np.int64)
    lib().orc_rot26(gp(geom), _p(np.ascontiguousarray(slice_index, np.int64)), k,
                    int(inverse), i, j, _p(io), _p(jo))
    return int(io[0]), int(jo[0])


def quadratic(geom, q, alpha, beta, lam, x):
    return lib().orc_quadratic(gp(geom), q, alpha, beta, lam,
                               _p(np.ascontiguousarray(x, np.float64)))


def edge_preserving(geom, q, alpha, beta, lam, delta, x):
    return lib().orc_edge_preserving(gp(geom), q, alpha, beta, lam, delta,
                                     _p(np.ascontiguousarray(x, np.float64)))


def rotate_slices(geom, tube_angles, slice_index, volume, direction):
    v = np.ascontiguousarray(volume, np.float32).copy()
    lib().orc_rotate_slices(gp(geom), _p(np.ascontiguousarray(tube_angles, np.float64)),
                            _p(np.ascontiguousarray(slice_index, np.int64)), _p(v),
                            direction)
    return v


def column_overlaps(geom, ssm, c, dv):
    cp, idx, _ = ssm
    return bool(lib().orc_column_overlaps(gp(geom), _p(cp), _p(idx), c, dv))
