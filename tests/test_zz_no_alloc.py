"""ctbb_cost_noalloc (allocation-free cost; no reference counterpart) against
ctbb_cost through the C ABI: device inputs, the result written to device
memory on a given stream, scratch from the library's static device array.
Two calls in a row, the second on a smaller input; an oversized geometry is
rejected; a matrix without its ray-major copy is refused.  Device buffers
and the stream come from the HIP runtime the library itself links."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip(ct):
    ct.init(0)
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    rt.hipStreamSynchronize.argtypes = [C.c_void_p]
    rt.hipStreamDestroy.argtypes = [C.c_void_p]
    return rt


class Dev:
    """A device copy of a numpy array (freed explicitly)."""

    def __init__(self, rt, a):
        self.rt = rt
        self.a = np.ascontiguousarray(a)
        self.p = C.c_void_p()
        assert rt.hipMalloc(C.byref(self.p), max(self.a.nbytes, 8)) == 0
        assert rt.hipMemcpy(self.p, self.a.ctypes.data_as(C.c_void_p), self.a.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.rt.hipFree(self.p)
            self.p = C.c_void_p()


def test_noalloc_cost_matches_ctbb_cost(ct, make_scan, hip):
    L = ct.lib()
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    for i, name in enumerate(["config0", "tiny_ffs"]):   # second call: smaller
        s = make_scan(name)
        g = s.geom
        raw, w = s.synthesize(use_gpu=True)
        vol = (s.phantom() * 0.9).astype(np.float32)
        opts = s.opts()
        m = ct.SSM.build(g)
        bufs = [Dev(hip, raw), None if w is None else Dev(hip, w),
                Dev(hip, s.slice_index.astype(np.int64)), Dev(hip, vol),
                Dev(hip, np.full(1, -1.0))]
        d_raw, d_w, d_si, d_vol, d_cost = bufs
        args = (m.h, d_raw.p, None if d_w is None else d_w.p, d_si.p, C.byref(opts),
                d_vol.p, d_cost.p, stream)
        print(f"[noalloc] {name}: {g.nx}x{g.ny}x{g.nz}, {g.readings} views", flush=True)
        if i == 0:
            # no ray-major copy yet: refused, nothing written
            assert L.ctbb_cost_noalloc(*args) == -6
            assert b"ray-major" in L.ctbb_last_error()
            assert d_cost.get()[0] == -1.0
        c_ref = ct.cost(m, raw, w, s.slice_index, opts, vol)   # builds the copy
        for rep in range(2):
            rc = L.ctbb_cost_noalloc(*args)
            assert rc == 0, L.ctbb_last_error()
            assert hip.hipStreamSynchronize(stream) == 0
            c = float(d_cost.get()[0])
            print(f"[noalloc] {name} call {rep}: {c:.12e} vs ctbb_cost {c_ref:.12e}",
                  flush=True)
            assert c == c_ref
        for b in bufs:
            if b is not None:
                b.free()
        m.free()
    hip.hipStreamDestroy(stream)


def test_noalloc_cost_rejects_oversized_geometry(ct, make_scan, hip):
    s = make_scan("tiny_ffs")
    g = ct.Geom.from_buffer_copy(s.geom)
    g.readings = 2_000_000          # 4e6 projection workgroups: beyond 8 MiB
    m = ct.SSM.build(g)
    d = Dev(hip, np.zeros(16, np.float32))
    cost = Dev(hip, np.zeros(1))
    L = ct.lib()
    opts = s.opts()
    rc = L.ctbb_cost_noalloc(m.h, d.p, None, d.p, C.byref(opts), d.p, cost.p, None)
    assert rc == -1, rc
    assert b"scratch" in L.ctbb_last_error()
    d.free()
    cost.free()
    m.free()
